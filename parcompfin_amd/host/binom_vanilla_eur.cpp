// binom_vanilla_eur -- drop-in for the reference program src/binom_vanilla_eur.cpp (main at :42-80).
//   binom_vanilla_eur call|put S0 E r sigma T N [flags]
// European option by backward induction on the beta tree, FP64, equal
// to the reference program's result bit for bit.
#include "cli_common.hpp"

int main(int argc, char** argv) {
    auto start_overall = cli::clock_t_::now();
    cli::Args a = cli::parse(argc, argv, 7, "binom_vanilla_eur call|put S0 E r sigma T N [flags]");
    pcf_ctx* ctx = cli::open_ctx();
    pcf_opts o = cli::base_opts(a);
    o.precision = PCF_FP64;
    pcf_result r{};
    auto start = cli::clock_t_::now();
    cli::check(ctx, pcf_binom_vanilla(ctx, a.S0, a.E, a.r, a.sigma, a.T, a.N, a.pf, 0, &o, &r),
               "binom_vanilla_eur");
    auto end = cli::clock_t_::now();
    std::chrono::duration<double> calc = end - start, overall = end - start_overall;
    cli::report(a, "HIP_vanilla", overall.count(), calc.count(), r.price, 0, 1);
    cli::stats_line(a, r);
    pcf_ctx_destroy(ctx);
    return 0;
}
