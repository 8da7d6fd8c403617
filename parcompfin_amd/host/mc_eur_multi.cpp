// mc_eur_multi -- drop-in for the reference program src/mc_eur_multi.cpp (main at :37-79).
//   mc_eur_multi call|put S0 E r sigma T N assets rho [flags]
// Extensions (optional):
//   --sigma-range LO,HI   per-asset volatilities uniform in [LO,HI] from --asset-seed
//   --random-corr SEED    correlation C = D^-1/2 (A A^T + d I) D^-1/2, A Gaussian d x d
//   --asset-seed SEED     seed of the per-asset draws (default 1)
// --rng mt19937 feeds standard normals from the host (verification form); the
// reference draws them with Boost's mt19937, which is not reproduced.
#include "cli_common.hpp"

int main(int argc, char** argv) {
    auto start_overall = cli::clock_t_::now();
    cli::Args a = cli::parse(argc, argv, 9, "mc_eur_multi call|put S0 E r sigma T N assets rho [flags]");
    const int d = a.assets;
    if (d < 1 || d > 64) cli::die("assets must be in [1, 64]");
    pcf_ctx* ctx = cli::open_ctx();
    pcf_opts o = cli::base_opts(a);
    unsigned long long aseed = 1;
    if (const char* s = cli::extra(a, "--asset-seed")) aseed = (unsigned long long)cli::inum(s, "--asset-seed");
    std::mt19937_64 g(aseed);
    std::vector<double> sig, corr;
    pcf_basket_ext ext{};
    bool use_ext = false;
    if (const char* s = cli::extra(a, "--sigma-range")) {
        double lo = 0, hi = 0;
        if (std::sscanf(s, "%lf,%lf", &lo, &hi) != 2 || !(hi >= lo) || lo < 0) cli::die("--sigma-range LO,HI");
        std::uniform_real_distribution<double> U(lo, hi);
        sig.resize(d);
        for (auto& v : sig) v = U(g);
        ext.sigma = sig.data();
        use_ext = true;
    }
    if (const char* s = cli::extra(a, "--random-corr")) {
        std::mt19937_64 gc((unsigned long long)cli::inum(s, "--random-corr"));
        std::normal_distribution<double> Z(0.0, 1.0);
        std::vector<double> A((size_t)d * d), C((size_t)d * d);
        for (auto& v : A) v = Z(gc);
        for (int i = 0; i < d; ++i)
            for (int k = 0; k < d; ++k) {
                double t = 0;
                for (int j = 0; j < d; ++j) t += A[i * d + j] * A[k * d + j];
                C[i * d + k] = t + (i == k ? d : 0.0);
            }
        corr.resize((size_t)d * d);
        for (int i = 0; i < d; ++i)
            for (int k = 0; k < d; ++k) corr[i * d + k] = C[i * d + k] / std::sqrt(C[i * d + i] * C[k * d + k]);
        ext.corr = corr.data();
        use_ext = true;
    }
    std::vector<double> z;
    if (a.mt19937) {
        z = cli::mt_normals(a.seed, 1.0, a.N * (long long)d);
        o.rng = PCF_RNG_HOST;
        o.host_normals = z.data();
        o.n_host_normals = (int64_t)z.size();
        o.ref_exact = 1;
    }
    pcf_result r{};
    auto start = cli::clock_t_::now();
    cli::check(ctx, pcf_mc_eur_multi(ctx, a.S0, a.E, a.r, a.sigma, a.T, a.N, a.pf, d, a.rho,
                                     use_ext ? &ext : nullptr, &o, &r), "mc_eur_multi");
    auto end = cli::clock_t_::now();
    std::chrono::duration<double> calc = end - start, overall = end - start_overall;
    cli::report(a, cli::method_name(a), overall.count(), calc.count(), r.price, 0, d);
    cli::stats_line(a, r);
    pcf_ctx_destroy(ctx);
    return 0;
}
