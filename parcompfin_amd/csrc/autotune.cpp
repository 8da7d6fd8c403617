// autotune.cpp -- launch-shape autotuner.
//
// A process-wide table maps (kind, arithmetic mode, size class) to a launch
// shape (threads per block, work items per thread).  Entry points consult it
// whenever their opts leave block_size / items_per_thread at 0, so every
// context -- including each serving slot -- launches a request of a given
// size with the same shape, and therefore reduces it with the same tree.
//
// pcf_autotune times candidate shapes through the public entry points on
// representative sizes (2^14, 2^16, ..., 2^max) and installs, for each class,
// the fastest one if it beats the built-in default by more than 3%; classes
// between measured sizes take the nearest measured entry.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pcf_internal.hpp"

namespace {

struct Shape {
    int B = 0, ipt = 0;
};
constexpr int kKinds = 5, kModes = 8, kClasses = 25;
std::mutex& tmu() {
    static std::mutex m;
    return m;
}
Shape (&table())[kKinds][kModes][kClasses] {
    static Shape t[kKinds][kModes][kClasses];
    return t;
}

struct Candidate {
    int B, ipt;
};

// Times one request shape: best kernel time of `reps` runs (after one warm-up).
double time_shape(pcf_ctx* ctx, int kind, int64_t N, int assets, int B, int ipt, int reps) {
    pcf_opts o;
    pcf_opts_init(&o);
    o.seed = 12345;
    o.precision = PCF_FP32K;
    o.block_size = B;
    o.items_per_thread = ipt;
    double best = 1e30;
    for (int r = 0; r <= reps; ++r) {
        pcf_result res{};
        int st;
        switch (kind) {
            case PCF_KIND_EUR:
                st = pcf_mc_eur(ctx, 100, 100, 0.05, 0.2, 1.0, N, 1.0, &o, &res);
                break;
            case PCF_KIND_ASIA:
                st = pcf_mc_asia(ctx, 100, 100, 0.05, 0.25, 1.0, N, 252, 1.0, &o, &res);
                break;
            case PCF_KIND_BASKET:
                st = pcf_mc_eur_multi(ctx, 100, 100, 0.05, 0.25, 1.0, N, 1.0, assets, 0.3, nullptr, &o, &res);
                break;
            default:
                o.kernel_variant = 2;  // the tuned shape is the persistent form's
                st = pcf_mc_amer(ctx, 100, 100, 0.06, 0.2, 1.0, N & ~(int64_t)1, 50, -1.0, &o, &res);
                break;
        }
        if (st != PCF_OK) return -1.0;
        if (r > 0) best = std::min(best, res.kernel_ms);
    }
    return best;
}

}  // namespace

void pcf_tuned_shape(int kind, int mode, int64_t N, int* B, int* ipt) {
    if (kind < 0 || kind >= kKinds || mode < 0 || mode >= kModes) return;
    const int c = pcf_size_class(N);
    std::lock_guard<std::mutex> lk(tmu());
    const Shape& s = table()[kind][mode][c];
    if (s.B) {
        *B = s.B;
        *ipt = s.ipt;
    }
}

extern "C" void pcf_tune_reset(void) {
    std::lock_guard<std::mutex> lk(tmu());
    for (auto& a : table())
        for (auto& b : a)
            for (auto& s : b) s = Shape{};
}

extern "C" int pcf_autotune(pcf_ctx* ctx, int32_t kind_mask, int32_t max_log2_paths, pcf_tune_entry* out,
                            int32_t max_out, int32_t* n_out) {
    if (!ctx) return PCF_ERR_INVALID_ARG;
    if (n_out) *n_out = 0;
    const int maxc = std::max(14, std::min(24, (int)max_log2_paths));
    struct Target {
        int kind, mode, assets;
        std::vector<Candidate> cands;
        Candidate def;
    };
    std::vector<Target> targets;
    auto grid = [](std::initializer_list<int> Bs, std::initializer_list<int> ipts) {
        std::vector<Candidate> v;
        for (int b : Bs)
            for (int i : ipts) v.push_back({b, i});
        return v;
    };
    if (kind_mask & (1 << PCF_KIND_EUR))
        targets.push_back({PCF_KIND_EUR, 0, 1, grid({64, 128, 256, 512, 1024}, {1, 2, 4, 8, 16}), {256, 4}});
    if (kind_mask & (1 << PCF_KIND_ASIA))
        targets.push_back({PCF_KIND_ASIA, 0, 1, grid({64, 128, 256, 512, 1024}, {1, 2}), {256, 1}});
    if (kind_mask & (1 << PCF_KIND_BASKET)) {
        // basket mode code = 2 * arithmetic mode + (MFMA ? 1 : 0): d = 8 on VALU, d = 32 on MFMA
        // ({0, 0}: the built-in shape, which grows the items per thread with N)
        targets.push_back({PCF_KIND_BASKET, 0, 8, grid({64, 128, 256}, {1, 2, 4, 8, 16}), {0, 0}});
        targets.push_back({PCF_KIND_BASKET, 1, 32, grid({64, 128, 256}, {1, 2, 4, 8, 16, 32}), {0, 0}});
    }
    if (kind_mask & (1 << PCF_KIND_AMER))
        // persistent LSM: (block size, pairs per thread); {0, 0} = built-in choice
        targets.push_back({PCF_KIND_AMER, 0, 1, grid({64, 128, 256, 512, 1024}, {1, 2, 4, 8}), {0, 0}});

    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    int nout = 0;
    for (const Target& t : targets) {
        std::vector<int> measured;
        for (int c = 14; c <= maxc; c += 2) {
            const int64_t N = (int64_t)1 << c;
            const int reps = c <= 18 ? 4 : 2;
            {
                std::lock_guard<std::mutex> tl(tmu());
                table()[t.kind][t.mode][c] = Shape{};  // the default is measured untuned
            }
            const double def_ms = time_shape(ctx, t.kind, N, t.assets, t.def.B, t.def.ipt, reps);
            if (def_ms < 0 && t.kind == PCF_KIND_AMER) continue;  // beyond the persistent form
            if (def_ms < 0) return pcf_fail(ctx, PCF_ERR_HIP, "autotune: default shape failed: %s",
                                            ctx->last_error.c_str());
            Candidate best = t.def;
            double best_ms = def_ms;
            int ncand = 0;
            for (const Candidate& cd : t.cands) {
                const double ms = time_shape(ctx, t.kind, N, t.assets, cd.B, cd.ipt, reps);
                if (ms < 0) continue;
                ++ncand;
                if (ms < best_ms) {
                    best_ms = ms;
                    best = cd;
                }
            }
            if (best_ms > 0.97 * def_ms) {
                best = t.def;
                best_ms = def_ms;
            }
            {
                std::lock_guard<std::mutex> tl(tmu());
                table()[t.kind][t.mode][c] = Shape{best.B, best.ipt};
            }
            measured.push_back(c);
            if (out && nout < max_out)
                out[nout] = pcf_tune_entry{t.kind, t.mode, c, best.B, best.ipt, ncand, best_ms, def_ms};
            ++nout;
        }
        // Unmeasured classes take the nearest measured one (for the persistent
        // LSM shape only inside the measured range: beyond it the built-in
        // choice adapts to the size and the graph form takes over).  Above the
        // measured range a tuned shape keeps its block count: the items per
        // thread double with each class.
        if (measured.empty()) continue;
        std::lock_guard<std::mutex> tl(tmu());
        for (int c = 10; c < kClasses; ++c) {
            int nearest = measured.front();
            for (int m : measured)
                if (std::abs(m - c) < std::abs(nearest - c)) nearest = m;
            if (t.kind == PCF_KIND_AMER && (c < measured.front() || c > measured.back())) continue;
            if (std::find(measured.begin(), measured.end(), c) == measured.end()) {
                Shape s = table()[t.kind][t.mode][nearest];
                if (s.B && c > nearest) s.ipt = std::min(1024, s.ipt << (c - nearest));
                table()[t.kind][t.mode][c] = s;
            }
        }
    }
    if (n_out) *n_out = std::min(nout, max_out);
    return PCF_OK;
}
