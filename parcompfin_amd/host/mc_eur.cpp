// mc_eur -- drop-in for the reference program src/mc_eur.cpp (main at :29-65).
//   mc_eur call|put S0 E r sigma T N [flags]
#include "cli_common.hpp"

int main(int argc, char** argv) {
    auto start_overall = cli::clock_t_::now();
    cli::Args a = cli::parse(argc, argv, 7, "mc_eur call|put S0 E r sigma T N [flags]");
    pcf_ctx* ctx = cli::open_ctx();
    pcf_opts o = cli::base_opts(a);
    std::vector<double> z;
    if (a.mt19937) {
        if (a.N <= 0) cli::die("N must be positive");
        z = cli::mt_normals(a.seed, std::sqrt(a.T), a.N);
        o.rng = PCF_RNG_HOST;
        o.host_normals = z.data();
        o.n_host_normals = (int64_t)z.size();
    }
    pcf_result r{};
    auto start = cli::clock_t_::now();
    cli::check(ctx, pcf_mc_eur(ctx, a.S0, a.E, a.r, a.sigma, a.T, a.N, a.pf, &o, &r), "mc_eur");
    auto end = cli::clock_t_::now();
    std::chrono::duration<double> calc = end - start, overall = end - start_overall;
    cli::report(a, cli::method_name(a), overall.count(), calc.count(), r.price, 0, 1);
    cli::stats_line(a, r);
    pcf_ctx_destroy(ctx);
    return 0;
}
