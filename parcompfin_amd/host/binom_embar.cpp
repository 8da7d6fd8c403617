// binom_embar -- drop-in for the reference program src/binom_embar.cpp (main at :53-91).
//   binom_embar call|put S0 E r sigma T N [flags]
// Default: the reference's comb() chain replayed bit for bit (FP64);
// --production: FP64 lgamma coefficients.
#include "cli_common.hpp"

int main(int argc, char** argv) {
    auto start_overall = cli::clock_t_::now();
    cli::Args a = cli::parse(argc, argv, 7, "binom_embar call|put S0 E r sigma T N [flags]");
    pcf_ctx* ctx = cli::open_ctx();
    pcf_opts o = cli::base_opts(a);
    o.ref_exact = a.production ? 0 : 1;
    o.precision = PCF_FP64;
    pcf_result r{};
    auto start = cli::clock_t_::now();
    cli::check(ctx, pcf_binom_embar(ctx, a.S0, a.E, a.r, a.sigma, a.T, a.N, a.pf, &o, &r), "binom_embar");
    auto end = cli::clock_t_::now();
    std::chrono::duration<double> calc = end - start, overall = end - start_overall;
    cli::report(a, "HIP_fp64", overall.count(), calc.count(), r.price, 0, 1);
    cli::stats_line(a, r);
    pcf_ctx_destroy(ctx);
    return 0;
}
