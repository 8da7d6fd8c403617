// mc_amer -- drop-in for the reference program src/mc_amer.cpp (main at :116-156).
//   mc_amer call|put S0 E r sigma T N M [flags]
// Default: corrected Longstaff-Schwartz (exercise test on payoff(S)).
// --ref-exact (implied by --rng mt19937) reproduces the reference's arithmetic,
// including its payoff(S-E,E) exercise test (src/mc_amer.cpp:100).
// --basis K (2..6) regresses on {1, x, .., x^(K-1)}, x = S/E - 1, instead of the
// reference's three terms (production arithmetic only).
#include "cli_common.hpp"

int main(int argc, char** argv) {
    auto start_overall = cli::clock_t_::now();
    cli::Args a = cli::parse(argc, argv, 8, "mc_amer call|put S0 E r sigma T N M [flags]");
    pcf_ctx* ctx = cli::open_ctx();
    pcf_opts o = cli::base_opts(a);
    if (const char* s = cli::extra(a, "--basis")) o.lsm_basis = (int32_t)cli::inum(s, "--basis");
    std::vector<double> w;
    if (a.mt19937) {
        if (a.N <= 0 || a.M <= 0 || a.N % 2) cli::die("N needs to be divisible by 2 for finding paths");
        o.ref_exact = 1;
        // pathsfinder seeds with time + 100*parallel (include/common.h:192), parallel = 0.
        w = cli::mt_normals(a.seed, std::sqrt(a.T / a.M), (a.N / 2) * (long long)a.M);
        o.rng = PCF_RNG_HOST;
        o.host_normals = w.data();
        o.n_host_normals = (int64_t)w.size();
    }
    pcf_result r{};
    auto start = cli::clock_t_::now();
    cli::check(ctx, pcf_mc_amer(ctx, a.S0, a.E, a.r, a.sigma, a.T, a.N, a.M, a.pf, &o, &r), "mc_amer");
    auto end = cli::clock_t_::now();
    std::chrono::duration<double> calc = end - start, overall = end - start_overall;
    cli::report(a, cli::method_name(a), overall.count(), calc.count(), r.price, a.M, 1);
    cli::stats_line(a, r);
    pcf_ctx_destroy(ctx);
    return 0;
}
