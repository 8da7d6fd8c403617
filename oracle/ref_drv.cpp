// ref_drv.cpp -- calls ONE reference kernel function, compiled from the
// reference's own source file (main renamed away by -Dmain=pcf_ref_main), and
// prints its result with 17 significant digits.
//
// TEST INFRASTRUCTURE ONLY: it produces the bit-level answers that pin the
// oracle restatement (oracle/oracle.cpp).  Seed = PCF_FIXTIME (fixtime.c).
//
// usage: drv_<prog> call|put S0 E r sigma T N [M]
#include <cstdio>
#include <cstdlib>
#include <cstring>

#if defined(DRV_MC_EUR)      // src/mc_eur.cpp:5-14
double mc_eur(double, double, double, double, double, int, double);
#elif defined(DRV_BINOM)     // src/binom_embar.cpp:5-14, src/binom_vanilla_{eur,amer}.cpp:5-14
double binom(double, double, double, double, double, int, double);
#elif defined(DRV_MC_ASIA)   // src/mc_asia.cpp:5-15
double mc_asia(double, double, double, double, double, int, int, double);
#elif defined(DRV_MC_AMER)   // src/mc_amer.cpp:5-15
double mc_amer(double, double, double, double, double, int, int, double);
#endif

int main(int argc, char** argv) {
    if (argc < 8) { std::fprintf(stderr, "usage: %s call|put S0 E r sigma T N [M]\n", argv[0]); return 2; }
    double pf = std::strcmp(argv[1], "put") == 0 ? -1.0 : 1.0;
    double S0 = std::atof(argv[2]), E = std::atof(argv[3]), r = std::atof(argv[4]);
    double sigma = std::atof(argv[5]), T = std::atof(argv[6]);
    int N = std::atoi(argv[7]);
    int M = argc > 8 ? std::atoi(argv[8]) : 0;
    (void)M;
    double v;
#if defined(DRV_MC_EUR)
    v = mc_eur(S0, E, r, sigma, T, N, pf);
#elif defined(DRV_BINOM)
    v = binom(S0, E, r, sigma, T, N, pf);
#elif defined(DRV_MC_ASIA)
    v = mc_asia(S0, E, r, sigma, T, N, M, pf);
#elif defined(DRV_MC_AMER)
    v = mc_amer(S0, E, r, sigma, T, N, M, pf);
#endif
    std::printf("%.17g\n", v);
    return 0;
}
