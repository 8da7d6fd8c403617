// cli_common.hpp -- shared pieces of the drop-in command-line programs.
//
// Each program keeps the reference's positional arguments
//     <prog> call|put S0 E r sigma T N [M] [assets rho]   (src/mc_eur.cpp:31-37,
//     src/mc_asia.cpp:51, src/mc_amer.cpp:125, src/mc_eur_multi.cpp:47-48)
// and prints the reference's one CSV row (include/common.h:212-249):
//     Method,Payoff,S0,E,r,sigma,T,N,M,Parallel,Nr_of_assets,T_overall,
//     T_calculation,Result,Abs_Error,Error
// with 10 significant digits.  Method is HIP (FP32/Kahan), HIP_fp64 or
// HIP_verify (--rng mt19937), and Parallel is the number of GPUs (1).  Everything else is an optional
// trailing flag, so runscript-style invocations keep working:
//     --seed S          Philox key / mt19937 seed (default: time(), as the reference)
//     --precision P     fp32k (default) | fp64
//     --rng mt19937     draw the reference's own libstdc++ mt19937 normal stream on
//                       the host and replay the reference arithmetic in FP64 on the
//                       device (reproduces the reference program's result)
//     --ref-exact       reference arithmetic choices (binom comb chain -- default for
//                       binom_embar; LSM quirks; basket without sqrt(T))
//     --production      binom_embar: double-double log-space coefficients (within
//                       ~1e-15 of the exact sum) instead of the reference's chain
//     --comparison X    value subtracted in the Abs_Error/Error columns (the
//                       reference compiles it in via include/comparison.h; default 0;
//                       the PCF_COMPARISON environment variable also works)
//     --comparison-file H   read it from a header of that form instead (the line
//                       "double comparison = X;" the runscripts write before they
//                       rebuild, runscript_mc_eur.sh:10-11); also PCF_COMPARISON_FILE
//     --header          print the runscripts' CSV header line before the row
//                       (runscript_mc_eur.sh:13)
//     --stats           one extra line on stderr: SE, kernel time, control values and,
//                       for an FP32/Kahan run, |FP32/Kahan - FP64| on the same Philox
//                       paths (the FP64 run is made for it)
// A trailing plain integer (the _omp programs' thread count) is accepted and ignored.
// Errors: message on stderr, exit status 1 (the reference aborts with an exception).
#pragma once

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <iomanip>
#include <iostream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "pcf.h"

namespace cli {

struct Args {
    std::string payoff;
    double pf = 1.0;
    double S0 = 0, E = 0, r = 0, sigma = 0, T = 0;
    long long N = 0;
    int M = 0;
    int assets = 1;
    double rho = 0;
    // flags
    unsigned long long seed = 0;
    bool seed_set = false;
    int precision = PCF_FP32K;
    bool mt19937 = false;
    bool ref_exact = false;
    bool production = false;
    double comparison = 0;
    bool header = false;
    bool stats = false;
    std::vector<std::string> extra;  // program-specific --key value pairs
};

[[noreturn]] inline void die(const std::string& msg) {
    std::fprintf(stderr, "error: %s\n", msg.c_str());
    std::exit(1);
}

// std::stod / std::stoi as the reference's getArgD / getArg (include/common.h:25-38).
inline double num(const char* s, const char* what) {
    try {
        return std::stod(s);
    } catch (...) {
        die(std::string("bad numeric argument for ") + what + ": '" + s + "'");
    }
}
inline long long inum(const char* s, const char* what) {
    try {
        std::size_t pos = 0;
        long long v = std::stoll(s, &pos);
        return v;
    } catch (...) {
        die(std::string("bad integer argument for ") + what + ": '" + s + "'");
    }
}

// The value of `comparison` in a header such as the runscripts write into
// include/comparison.h ("#pragma once" and "double comparison = X;",
// runscript_mc_eur.sh:10-11).  Comments are not parsed; the first
// "comparison" followed by '=' gives the value.
inline double comparison_from_header(const char* path) {
    std::FILE* f = std::fopen(path, "r");
    if (!f) die(std::string("cannot read comparison header '") + path + "'");
    std::string text;
    char buf[4096];
    for (std::size_t n; (n = std::fread(buf, 1, sizeof buf, f)) > 0;) text.append(buf, n);
    std::fclose(f);
    for (std::size_t at = text.find("comparison"); at != std::string::npos; at = text.find("comparison", at + 1)) {
        std::size_t eq = text.find_first_not_of(" \t\r\n", at + 10);
        if (eq == std::string::npos || text[eq] != '=') continue;
        const std::size_t end = text.find(';', eq);
        if (end == std::string::npos) break;
        return num(text.substr(eq + 1, end - eq - 1).c_str(), "comparison in the header");
    }
    die(std::string("no 'comparison = value;' in '") + path + "'");
}

// npos: number of positional arguments after the program name (7, 8 or 9).
inline Args parse(int argc, char** argv, int npos, const char* usage) {
    Args a;
    if (argc < 1 + npos) die(std::string("usage: ") + usage);
    a.payoff = argv[1];
    if (a.payoff == "call") a.pf = 1.0;
    else if (a.payoff == "put") a.pf = -1.0;
    else die("Unknown payoff function");
    a.S0 = num(argv[2], "S0");
    a.E = num(argv[3], "E");
    a.r = num(argv[4], "r");
    a.sigma = num(argv[5], "sigma");
    a.T = num(argv[6], "T");
    a.N = inum(argv[7], "N");
    if (npos >= 8) {
        if (npos == 8) a.M = (int)inum(argv[8], "M");
        else {
            a.assets = (int)inum(argv[8], "assets");
            a.rho = num(argv[9], "rho");
        }
    }
    if (const char* c = std::getenv("PCF_COMPARISON_FILE")) a.comparison = comparison_from_header(c);
    if (const char* c = std::getenv("PCF_COMPARISON")) a.comparison = num(c, "PCF_COMPARISON");
    for (int i = 1 + npos; i < argc; ++i) {
        std::string f = argv[i];
        auto val = [&](const char* what) -> const char* {
            if (i + 1 >= argc) die(std::string("missing value for ") + what);
            return argv[++i];
        };
        if (f == "--seed") {
            a.seed = (unsigned long long)inum(val("--seed"), "--seed");
            a.seed_set = true;
        } else if (f == "--precision") {
            std::string p = val("--precision");
            if (p == "fp32k") a.precision = PCF_FP32K;
            else if (p == "fp64") a.precision = PCF_FP64;
            else die("--precision must be fp32k or fp64");
        } else if (f == "--rng") {
            std::string p = val("--rng");
            if (p == "mt19937") a.mt19937 = true;
            else if (p != "philox") die("--rng must be philox or mt19937");
        } else if (f == "--ref-exact") {
            a.ref_exact = true;
        } else if (f == "--production") {
            a.production = true;
        } else if (f == "--comparison") {
            a.comparison = num(val("--comparison"), "--comparison");
        } else if (f == "--comparison-file") {
            a.comparison = comparison_from_header(val("--comparison-file"));
        } else if (f == "--header") {
            a.header = true;
        } else if (f == "--stats") {
            a.stats = true;
        } else if (f == "--control") {
            a.extra.push_back(f);
            a.extra.push_back("1");
        } else if (f.rfind("--", 0) == 0) {
            a.extra.push_back(f);
            a.extra.push_back(val(f.c_str()));
        } else {
            (void)inum(argv[i], "threads");  // the _omp programs' trailing thread count
        }
    }
    if (!a.seed_set) a.seed = (unsigned long long)std::time(nullptr);
    return a;
}

inline const char* extra(const Args& a, const char* key) {
    for (size_t i = 0; i + 1 < a.extra.size(); i += 2)
        if (a.extra[i] == key) return a.extra[i + 1].c_str();
    return nullptr;
}

// The reference's normal stream: libstdc++ mt19937 seeded with `seed`,
// normal_distribution<>{0, sd} (src/mc_eur.cpp:16-20).
inline std::vector<double> mt_normals(unsigned long long seed, double sd, long long count) {
    std::mt19937 gen;
    gen.seed(seed);
    std::normal_distribution<> norm{0, sd};
    std::vector<double> z((size_t)count);
    for (auto& v : z) v = norm(gen);
    return z;
}

inline pcf_ctx* open_ctx() {
    pcf_ctx* ctx = nullptr;
    int st = pcf_ctx_create(0, &ctx);
    if (st != PCF_OK) die("no usable HIP device (pcf_ctx_create status " + std::to_string(st) + ")");
    return ctx;
}

inline void check(pcf_ctx* ctx, int st, const char* what) {
    if (st != PCF_OK) die(std::string(what) + ": " + pcf_ctx_last_error(ctx));
}

using clock_t_ = std::chrono::system_clock;

// One row, formatted as include/common.h:231-248 does.
inline void report(const Args& a, const char* method, double t_overall, double t_calc, double result,
                   int M, int assets) {
    if (a.header)
        std::cout << "Method,Payoff,S0,E,r,sigma,T,N,M,Parallel,Nr_of_assets,T_overall,T_calculation,Result,"
                     "Abs_Error,Error\n";
    std::cout << std::setprecision(10) << method << "," << a.payoff << "," << a.S0 << "," << a.E << ","
              << a.r << "," << a.sigma << "," << a.T << "," << a.N << "," << M << "," << 1 << ","
              << assets << "," << t_overall << "," << t_calc << "," << result << ","
              << std::fabs(result - a.comparison) << "," << result - a.comparison << std::endl;
}

inline const char* method_name(const Args& a) {
    if (a.mt19937) return "HIP_verify";
    return a.precision == PCF_FP64 ? "HIP_fp64" : "HIP";
}

inline void stats_line(const Args& a, const pcf_result& r) {
    if (!a.stats) return;
    std::fprintf(stderr,
                 "stats: price=%.17g se=%.6g kernel_ms=%.6f total_ms=%.6f control=%.10g control_se=%.3g "
                 "paths=%lld steps=%lld small_itm=%d fallback=%d grid=%d block=%d seed=%llu",
                 r.price, r.std_err, r.kernel_ms, r.total_ms, r.control_price, r.control_se,
                 (long long)r.n_paths, (long long)r.n_steps, r.lsm_small_itm, r.lsm_fallback,
                 r.grid_blocks, r.block_size, a.seed);
    if (r.fp64_checked)
        std::fprintf(stderr, " fp64_err=%.6g fp64_err_se=%.4g", r.fp64_err,
                     r.std_err > 0 ? r.fp64_err / r.std_err : 0.0);
    std::fprintf(stderr, "\n");
}

inline pcf_opts base_opts(const Args& a) {
    pcf_opts o;
    pcf_opts_init(&o);
    o.seed = a.seed;
    o.precision = a.precision;
    o.ref_exact = a.ref_exact ? 1 : 0;
    // --stats on an FP32/Kahan run also prices the same paths in FP64 and
    // reports the difference
    o.fp64_check = a.stats ? 1 : 0;
    return o;
}

}  // namespace cli
