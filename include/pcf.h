/* pcf.h -- C-ABI of libpcf, the MI355X (gfx950) Monte Carlo / binomial
 * pricing engine.
 *
 * Plain C types only: no HIP, no torch, no C++ across this boundary.  Every
 * entry point returns an int status (PCF_OK = 0); no exception crosses it.
 * Caller owns all parameter and result memory; a pcf_ctx owns every device
 * buffer (resident between calls) and one HIP stream.  A context is used by
 * one host thread at a time (calls on one context are serialised internally).
 *
 * Each pricing entry point replaces one reference kernel function and keeps
 * its argument list and meaning (S0, E, r, sigma, T, N[, M][, assets, rho],
 * payoff_fun = +1 call / -1 put), then adds options and a result record.
 */
#ifndef PCF_H
#define PCF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCF_ABI_VERSION 2

/* ---- status codes ------------------------------------------------------ */
enum pcf_status {
    PCF_OK = 0,
    PCF_ERR_INVALID_ARG = 1,   /* bad payoff / size / parameter (the reference throws
                                  std::invalid_argument: src/mc_eur.cpp:42,
                                  include/common.h:180) */
    PCF_ERR_NO_DEVICE = 2,     /* no HIP device / device index out of range */
    PCF_ERR_HIP = 3,           /* a HIP runtime call failed (message in last_error) */
    PCF_ERR_OUT_OF_MEMORY = 4, /* device allocation failed, or a request exceeds the
                                  server's memory budget */
    PCF_ERR_SINGULAR = 5,      /* reference-exact LSM: Gram determinant <= 0
                                  (include/common.h:115-117 throws) */
    PCF_ERR_NOT_FOUND = 6,     /* unknown server ticket */
    PCF_ERR_SHUTDOWN = 7,      /* server is shutting down */
    PCF_ERR_TIMEOUT = 8        /* a bounded device wait expired */
};

/* ---- arithmetic form ---------------------------------------------------- */
enum pcf_precision {
    PCF_FP32K = 0, /* FP32 path evolution and payoffs; per-thread and per-block payoff
                      sums Kahan-compensated in FP64; FP64 final mean/variance */
    PCF_FP64 = 1   /* everything in FP64 */
};

/* ---- normal source ------------------------------------------------------ */
enum pcf_rng {
    PCF_RNG_PHILOX = 0, /* counter-based Philox4x32-10 + Box-Muller, keyed by
                           (seed, path, step): key = (seed lo, seed hi),
                           counter = (step/4, path lo, path hi, stream id);
                           uniform u = (x + 0.5) 2^-32; the pair (u0,u1) gives
                           R cos(2 pi u1), R sin(2 pi u1), R = sqrt(-2 ln u0).
                           Any path is reproducible regardless of launch shape. */
    PCF_RNG_HOST = 1    /* verification: the caller passes FP64 normals in the
                           reference's draw order (see each entry point); the
                           device replays the reference arithmetic in FP64 */
};

/* Philox stream ids (counter word 3). */
enum pcf_stream {
    PCF_STREAM_EUR = 1,    /* normal p%4 of block 0 of Philox path p/4 */
    PCF_STREAM_ASIA = 2,   /* step m of path p: normal m of path p */
    PCF_STREAM_AMER = 3,   /* date m (1..M) of antithetic pair n: normal m-1 of path n */
    PCF_STREAM_BASKET = 4  /* asset a of path p: normal a of path p */
};

typedef struct pcf_opts {
    uint64_t seed;           /* Philox key */
    int32_t precision;       /* enum pcf_precision */
    int32_t rng;             /* enum pcf_rng */
    const double* host_normals; /* PCF_RNG_HOST only */
    int64_t n_host_normals;
    int32_t ref_exact;       /* 1: replay the reference's arithmetic choices exactly
                                (binom: comb() chain; mc_amer: basis {1,S-E,(S-E)^2},
                                cofactor inverse and the payoff(S-E,E) exercise test;
                                basket: no sqrt(T) on the correlated draws) */
    int32_t control;         /* mc_asia: also price the geometric-average control */
    int32_t block_size;      /* threads per block; 0 = tuned default */
    int32_t items_per_thread;/* paths (or path groups) per thread; 0 = tuned default */
    int32_t* when_out;       /* mc_amer: optional host array of N exercise dates */
    int32_t kernel_variant;  /* 0 = automatic; mc_eur_multi FP32: 1 = VALU FMA kernel,
                                2 = MFMA kernel (needs 16 <= assets <= 32);
                                mc_amer: 1 = graph of per-date launches, 2 = one
                                persistent launch (error if its grid cannot be
                                co-resident; block_size / items_per_thread, when
                                set, give its block size and pairs per thread) */
    int32_t lsm_basis;       /* mc_amer: regression basis size K; 0 = 3 ({1, S/E, (S/E)^2}, or the
                                reference's {1, S-E, (S-E)^2} with ref_exact); 2..6 = {1, x, ..,
                                x^(K-1)}, x = S/E - 1 (production arithmetic; persistent and
                                graph forms) */
    int32_t fp64_check;      /* 1: a PCF_FP32K Philox pricing (mc_eur, mc_asia, mc_amer,
                                mc_eur_multi, direct or served) is repeated in PCF_FP64 on the
                                same Philox counters, and pcf_result.fp64_err receives
                                |price - FP64 price|; the timing fields stay the FP32 run's */
    int32_t reserved[3];
} pcf_opts;

typedef struct pcf_result {
    double price;        /* discounted price: the reference function's return value */
    double std_err;      /* standard error of price (0 for the binomial sum) */
    double mean;         /* FP64 mean of undiscounted path payoffs (binom: term sum) */
    double variance;     /* FP64 sample variance of undiscounted path payoffs */
    double control_price;/* mc_asia: geometric-average control price on the same paths;
                            mc_amer: European price on the same paths (no exercise) */
    double control_se;
    int64_t n_paths;
    int64_t n_steps;     /* time steps per path (1 for single-step products) */
    double kernel_ms;    /* device time of the dominant kernel, HIP events on the
                            launch stream */
    double total_ms;     /* device time of the whole launch sequence */
    double flops;        /* algorithmic flop count of the dominant kernel */
    double bytes;        /* algorithmic HBM bytes of the dominant kernel */
    int32_t lsm_small_itm; /* mc_amer: dates with 1..2 in-the-money paths */
    int32_t lsm_fallback;  /* mc_amer: dates whose Gram solve fell back to a smaller basis;
                              mc_eur_multi: 1 if the correlation factor came from the
                              eigen-decomposition (not positive definite) */
    int32_t grid_blocks;   /* launch shape of the dominant kernel */
    int32_t block_size;
    int32_t batch_size;    /* requests priced by the same launch (1 when alone) */
    int32_t fp64_checked;  /* 1 when fp64_err holds the FP64 comparison (opts.fp64_check) */
    double fp64_err;       /* |price - price of the FP64 form on the same Philox paths| */
} pcf_result;

typedef struct pcf_ctx pcf_ctx;

/* ---- context ------------------------------------------------------------ */
/* A context owns one HIP stream and its resident buffers; use it from one
 * thread at a time.  Contexts of one process on one device share the
 * ordering of mc_amer's persistent launches: each runs after the kernels the
 * other contexts had enqueued, and the other contexts' heavy launches
 * (about 40 us of work and up) run after the last persistent launch. */
int pcf_ctx_create(int device, pcf_ctx** out);
int pcf_ctx_destroy(pcf_ctx* ctx);
const char* pcf_ctx_last_error(const pcf_ctx* ctx);
int pcf_ctx_synchronize(pcf_ctx* ctx);
/* Device name, compute units, device memory in bytes. */
int pcf_device_info(pcf_ctx* ctx, char* name, int name_len, int* compute_units,
                    int64_t* total_mem);
void pcf_opts_init(pcf_opts* opts);
int pcf_abi_version(void);
/* sizeof of an ABI record by name ("pcf_opts", "pcf_result", "pcf_basket_ext",
 * "pcf_request", "pcf_server_config", "pcf_request_timing", "pcf_tune_entry",
 * "pcf_server_stats");
 * -1 if unknown.
 * Lets FFI bindings verify their struct mirrors without a device. */
int pcf_struct_size(const char* name);

/* ---- pricing entry points ----------------------------------------------- */

/* European option on one GBM asset, single terminal step.
 * Replaces double mc_eur(S0,E,r,sigma,T,N,payoff_fun), src/mc_eur.cpp:5-27.
 * PCF_RNG_HOST: host_normals[n], n < N, are the draws N(0, sqrt(T)). */
int pcf_mc_eur(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
               int64_t N, double payoff_fun, const pcf_opts* opts, pcf_result* res);

/* Closed binomial sum over the beta tree (src/binom_embar.cpp:19-27).
 * Replaces double binom(S0,E,r,sigma,T,N,payoff_fun), src/binom_embar.cpp:5-50.
 * ref_exact = 1 (default through the CLI) replays comb() of include/common.h:63-72
 * bit for bit; ref_exact = 0 forms each weight C(N,i) p^i q^(N-i) in log space
 * in double-double arithmetic (a resident table of log k!), within ~1e-15 of
 * the exact sum (the reference's chain is 1.2e-10 from it at N = 1e5 and
 * 1.5e-8 at N = 1e6). */
int pcf_binom_embar(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
                    int64_t N, double payoff_fun, const pcf_opts* opts, pcf_result* res);

/* Fixed-strike arithmetic Asian with the reference's trapezoid running sum.
 * Replaces double mc_asia(S0,E,r,sigma,T,N,M,payoff_fun), src/mc_asia.cpp:5-40.
 * PCF_RNG_HOST: host_normals[n*M+m] are the draws N(0, sqrt(T/M)); with
 * kernel_variant 0 the device replays the reference's FP64 arithmetic on them,
 * with kernel_variant 1 it runs the production FP64 step (I += S (a1 + a2 z),
 * S *= exp(mu + vol z)) on z = draw / sqrt(T/M), the arithmetic of the Philox
 * forms on the reference's stream. */
int pcf_mc_asia(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
                int64_t N, int32_t M, double payoff_fun, const pcf_opts* opts,
                pcf_result* res);

/* pcf_mc_asia without device allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory -- tight serving loops and
 * harnesses that forbid hipMalloc.  It runs pcf_mc_asia's kernels with the
 * same launch shape, so its result equals pcf_mc_asia's bit for bit, but every
 * device buffer those kernels use (job record, per-block partials, arrival
 * counters, results) lives in one file-scope device array of 2 MiB + 5 KiB
 * per GPU, held until the process ends.  Limits: Philox normals only (host
 * normals would need a device buffer), no fp64_check, and at most 32768 grid
 * blocks with the geometric control (65536 without); larger launches are
 * refused with PCF_ERR_INVALID_ARG.  The context's stream, events and pinned
 * staging (allocated by pcf_ctx_create) are used; the call allocates nothing.
 * Calls on one GPU share the array and are serialised by a per-GPU lock (each
 * call returns after its work completes); the call cannot be captured into a
 * graph, as it synchronises its stream. */
int pcf_mc_asia_noalloc(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
                        int64_t N, int32_t M, double payoff_fun, const pcf_opts* opts,
                        pcf_result* res);

/* American option by Longstaff-Schwartz on an antithetic path store.
 * Replaces double mc_amer(S0,E,r,sigma,T,N,M,payoff_fun), src/mc_amer.cpp:5-114
 * (paths: include/common.h:168-208).  N must be even.
 * PCF_RNG_HOST: host_normals[n*M+(m-1)], n < N/2, m = 1..M, are the draws
 * N(0, sqrt(T/M)) in pathsfinder's order. */
int pcf_mc_amer(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
                int64_t N, int32_t M, double payoff_fun, const pcf_opts* opts,
                pcf_result* res);

/* Backward induction on the same beta tree (the reference's lattice programs):
 * american = 0 replaces double binom(S0,E,r,sigma,T,N,payoff_fun) of
 * src/binom_vanilla_eur.cpp:5-39, american = 1 the one of
 * src/binom_vanilla_amer.cpp:5-40.  FP64, the reference's operations and
 * glibc powers, so the result equals the reference program's bit for bit.
 * N <= 2^21. */
int pcf_binom_vanilla(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
                      int64_t N, double payoff_fun, int32_t american, const pcf_opts* opts,
                      pcf_result* res);

/* Optional per-asset extension of the basket (NULL pointers = the reference's
 * scalar sigma, S0, weights 1/d and equicorrelation rho). */
typedef struct pcf_basket_ext {
    const double* sigma;   /* d volatilities */
    const double* S0;      /* d spot prices */
    const double* weights; /* d weights */
    const double* corr;    /* d x d correlation matrix, row-major, SPD */
} pcf_basket_ext;

/* Arithmetic basket call/put on d correlated GBM assets, single terminal step.
 * Replaces double mc_eur(S0,E,r,sigma,T,N,payoff_fun,assets,rho),
 * src/mc_eur_multi.cpp:6-35, with mvnorm of include/mvn.h:42-82.
 * PCF_RNG_HOST: host_normals[n*d+a] are standard normals (column-major d x N,
 * the order of mvnorm's NullaryExpr); the device applies the Cholesky factor. */
int pcf_mc_eur_multi(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
                     int64_t N, double payoff_fun, int32_t assets, double rho,
                     const pcf_basket_ext* ext, const pcf_opts* opts, pcf_result* res);

/* ---- autotuner --------------------------------------------------------- */
/* One tuned launch shape: threads per block and work items per thread for a
 * (kind, arithmetic mode, size class = ceil(log2 N)) triple, with the best
 * measured kernel time. */
typedef struct pcf_tune_entry {
    int32_t kind, mode, size_class, block_size, items_per_thread, candidates;
    double best_ms, default_ms;
} pcf_tune_entry;

/* Sweeps block size and items per thread for the Philox FP32/Kahan forms of
 * the kinds in kind_mask (bit k = enum pcf_kind k; EUR, ASIA, BASKET, AMER) at
 * size classes 2^14 .. 2^max_log2_paths (AMER: block size and antithetic
 * pairs per thread of the persistent form, at the sizes it holds), times each
 * candidate on the device,
 * and installs the fastest shape in the process-wide table every context
 * (and the server) consults when opts leave block_size / items_per_thread 0.
 * A shape change moves the reduction tree, so results change in the last
 * bits.  Fills up to max_out entries; *n_out receives the count. */
int pcf_autotune(pcf_ctx* ctx, int32_t kind_mask, int32_t max_log2_paths, pcf_tune_entry* out,
                 int32_t max_out, int32_t* n_out);
/* Forgets every tuned shape (back to the built-in defaults). */
void pcf_tune_reset(void);

/* Testing hook for the process-wide basket factor cache (no counterpart in
 * the reference): the cache finds an entry by a 64-bit hash of the basket's
 * parameters ANDed with `mask` (default all ones) and uses it only when the
 * entry's full parameter key equals the request's.  mask = 0 sends every
 * basket to one hash value, so tests can show that collisions never mix
 * factors. */
void pcf_debug_basket_hash_mask(uint64_t mask);

/* Standard normals of the device generator: out[p*K + k] = normal k of path
 * path0+p of `stream` (the exact values the pricing kernels consume, in FP64
 * or rounded to FP32 per `precision`).  For reproducibility checks. */
int pcf_normals(pcf_ctx* ctx, uint64_t seed, uint32_t stream, uint64_t path0,
                int64_t n_paths, int32_t K, int32_t precision, double* out);

/* ---- serving ------------------------------------------------------------ */
enum pcf_kind { PCF_KIND_EUR = 0, PCF_KIND_ASIA = 1, PCF_KIND_AMER = 2,
                PCF_KIND_BASKET = 3, PCF_KIND_BINOM = 4 };

typedef struct pcf_request {
    int32_t kind;        /* enum pcf_kind */
    int32_t M;           /* steps (asia) / exercise dates (amer) */
    int32_t assets;      /* basket dimension */
    int32_t pad;
    double S0, E, r, sigma, T, payoff_fun, rho;
    int64_t N;
    pcf_opts opts;       /* seed, precision; rng must be PCF_RNG_PHILOX */
    const pcf_basket_ext* basket; /* basket extension or NULL; must stay valid
                                     until pcf_server_wait returns */
} pcf_request;

typedef struct pcf_server_config {
    int64_t mem_budget_bytes; /* device bytes the server may hold for requests; 0 = 8 GiB */
    int32_t max_inflight;     /* concurrent requests on the device; 0 = 8 */
    int32_t use_graphs;       /* 1: replay captured hipGraphs per (kind, size class) */
    int32_t reserved[8];
} pcf_server_config;

typedef struct pcf_request_timing {
    double submit_ms;    /* host clock at submit, ms since server start */
    double start_ms;     /* admitted and launched */
    double done_ms;      /* device work complete */
    double latency_ms;   /* done - submit */
} pcf_request_timing;

typedef struct pcf_server pcf_server;

int pcf_server_create(pcf_ctx* ctx, const pcf_server_config* cfg, pcf_server** out);
/* Queue a request; returns at once with a ticket.  PCF_ERR_OUT_OF_MEMORY if the
 * request alone can never fit the budget. */
int pcf_server_submit(pcf_server* srv, const pcf_request* req, uint64_t* ticket);
/* Block until the ticket's request completes; the ticket is then released. */
int pcf_server_wait(pcf_server* srv, uint64_t ticket, pcf_result* res,
                    pcf_request_timing* timing);
/* Drains queued requests, joins the slots and frees everything. */
int pcf_server_destroy(pcf_server* srv);
/* Message of the last refused submit or failed launch. */
const char* pcf_server_last_error(pcf_server* srv);

/* Device memory of a server.  Each slot is charged for the request-sized
 * buffers it holds or may hold (path stores and per-path state, reduction
 * partials, binomial tables): what it kept from its last request when idle,
 * plus, when busy, what the unit it runs can add (for an American request
 * the growth of the kept path store and state it reuses, plus its other
 * buffers).  The sum of the charges never exceeds mem_budget_bytes: idle
 * slots' kept buffers (least recently used first), then the slot's own, are
 * freed before a request waits. */
typedef struct pcf_server_stats {
    int64_t budget_bytes;         /* mem_budget_bytes in force */
    int64_t charged_bytes;        /* sum of the slots' charges now */
    int64_t peak_charged_bytes;   /* largest charged_bytes so far (<= budget_bytes) */
    int64_t held_bytes;           /* request-sized bytes the slots held when they last went idle */
    int64_t peak_held_bytes;      /* largest held_bytes at a unit's completion */
    int64_t fixed_bytes_per_slot; /* per-slot buffers outside the budget (arrival counters,
                                     results, job table, basket factor slots) */
    int64_t evictions;            /* kept buffers freed to admit a request (a slot's own, or
                                     an idle slot's) */
    int64_t underestimates;       /* units that held more than their charge (then freed); 0 */
    int64_t completed;            /* requests completed */
    int32_t slots;
    int32_t reserved[5];
} pcf_server_stats;
int pcf_server_get_stats(pcf_server* srv, pcf_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* PCF_H */
