// mc_asia -- drop-in for the reference program src/mc_asia.cpp (main at :42-81).
//   mc_asia call|put S0 E r sigma T N M [flags]      (--control: geometric control on stderr)
#include "cli_common.hpp"

int main(int argc, char** argv) {
    auto start_overall = cli::clock_t_::now();
    cli::Args a = cli::parse(argc, argv, 8, "mc_asia call|put S0 E r sigma T N M [flags]");
    pcf_ctx* ctx = cli::open_ctx();
    pcf_opts o = cli::base_opts(a);
    o.control = cli::extra(a, "--control") ? 1 : 0;
    std::vector<double> z;
    if (a.mt19937) {
        if (a.N <= 0 || a.M <= 0) cli::die("N and M must be positive");
        z = cli::mt_normals(a.seed, std::sqrt((double)a.T / (double)a.M), a.N * (long long)a.M);
        o.rng = PCF_RNG_HOST;
        o.host_normals = z.data();
        o.n_host_normals = (int64_t)z.size();
    }
    pcf_result r{};
    auto start = cli::clock_t_::now();
    cli::check(ctx, pcf_mc_asia(ctx, a.S0, a.E, a.r, a.sigma, a.T, a.N, a.M, a.pf, &o, &r), "mc_asia");
    auto end = cli::clock_t_::now();
    std::chrono::duration<double> calc = end - start, overall = end - start_overall;
    cli::report(a, cli::method_name(a), overall.count(), calc.count(), r.price, a.M, 1);
    cli::stats_line(a, r);
    pcf_ctx_destroy(ctx);
    return 0;
}
