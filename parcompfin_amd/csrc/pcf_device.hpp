// pcf_device.hpp -- device building blocks shared by the gfx950 pricing
// kernels: counter-based Philox4x32-10, Box-Muller normals (FP32 on the
// hardware transcendental unit, FP64 on the libm sequences), payoff, and
// deterministic compensated FP64 reductions (wavefront shuffles + LDS, then a
// last-arriving-block pass over the per-block partials).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcf {

constexpr int kWave = 64;  // CDNA wavefront width

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC11).  Multipliers 0xD2511F53/0xCD9E8D57,
// Weyl increments 0x9E3779B9/0xBB67AE85.
struct U4 {
    uint32_t x, y, z, w;
};

// a ^ b ^ c as one gfx950 v_bitop3_b32 (truth table 0x96, the parity of the
// three input bits).  Written as two ^ the compiler issues two v_xor_b32,
// and the Monte Carlo kernels are bound by VALU issue, a third of it Philox.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// Each round needs both halves of two 32x32 products; written as one 64-bit
// product so gfx950 issues a single v_mad_u64_u32 per product (measured
// ~2.4x the throughput of a v_mul_lo_u32 + v_mul_hi_u32 pair).
__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        U4 n;
        if (round == 0) {
            // Counter words are often wave-uniform (block, stream): left as
            // plain xors, the compiler keeps this round on the scalar unit.
            n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
            n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        } else {
            n.x = xor3((uint32_t)(p1 >> 32), c.y, k0);
            n.z = xor3((uint32_t)(p0 >> 32), c.w, k1);
        }
        n.y = (uint32_t)p1;
        n.w = (uint32_t)p0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Counter layout: (block of 4 normals, path lo, path hi, stream id).
__device__ __forceinline__ U4 philox_block(uint64_t seed, uint32_t stream, uint64_t path,
                                           uint32_t block) {
    U4 c;
    c.x = block;
    c.y = (uint32_t)path;
    c.z = (uint32_t)(path >> 32);
    c.w = stream;
    return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// ---------------------------------------------------------------------------
// Box-Muller.  Uniform u = (x + 0.5) 2^-32 in (0,1); the pair (u0, u1) gives
// z0 = R cos(2 pi u1), z1 = R sin(2 pi u1), R = sqrt(-2 ln u0).
//
// FP32: v_log_f32 is log2, v_sin_f32 / v_cos_f32 take their argument in
// revolutions (sin(2 pi x)), so one instruction each; u is the FP32 rounding
// of the FP64 uniform.
__device__ __forceinline__ float u01_f32(uint32_t x) {
    return fmaf((float)x, 0x1p-32f, 0x1p-33f);
}

__device__ __forceinline__ void box_muller_f32(uint32_t a, uint32_t b, float& z0, float& z1) {
    const float u0 = u01_f32(a);
    const float u1 = u01_f32(b);
    // -2 ln u0 = -2 ln(2) log2(u0)
    const float R = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u0));
    z0 = R * __builtin_amdgcn_cosf(u1);
    z1 = R * __builtin_amdgcn_sinf(u1);
}

__device__ __forceinline__ double u01_f64(uint32_t x) {
    return ((double)x + 0.5) * 0x1p-32;
}

__device__ __forceinline__ void box_muller_f64(uint32_t a, uint32_t b, double& z0, double& z1) {
    const double u0 = u01_f64(a);
    const double u1 = u01_f64(b);
    const double R = sqrt(-2.0 * log(u0));
    double s, c;
    sincos(6.283185307179586476925286766559 * u1, &s, &c);
    z0 = R * c;
    z1 = R * s;
}

template <typename Real>
__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, Real& z0, Real& z1);
template <>
__device__ __forceinline__ void box_muller<float>(uint32_t a, uint32_t b, float& z0, float& z1) {
    box_muller_f32(a, b, z0, z1);
}
template <>
__device__ __forceinline__ void box_muller<double>(uint32_t a, uint32_t b, double& z0, double& z1) {
    box_muller_f64(a, b, z0, z1);
}

// Four standard normals of (seed, stream, path, block).
template <typename Real>
__device__ __forceinline__ void normals4(uint64_t seed, uint32_t stream, uint64_t path,
                                         uint32_t block, Real z[4]) {
    const U4 x = philox_block(seed, stream, path, block);
    box_muller<Real>(x.x, x.y, z[0], z[1]);
    box_muller<Real>(x.z, x.w, z[2], z[3]);
}

// ---------------------------------------------------------------------------
// Fast FP32 exp on the hardware exp2 unit: exp(x) = 2^(x log2 e).  Kernels
// fold log2(e) into their constants and call exp2 directly.
__device__ __forceinline__ float exp2_f32(float x) { return __builtin_amdgcn_exp2f(x); }

template <typename Real>
__device__ __forceinline__ Real payoff(Real S, Real E, Real pf) {
    const Real v = pf * (S - E);
    return v > Real(0) ? v : Real(0);
}

// ---------------------------------------------------------------------------
// Compensated FP64 sums.  value = s + c; TwoSum (Knuth) keeps the rounding
// error of every add in c, so a sum of sums is accurate to ~1 ulp of the
// total independent of the order of the partials.
struct CSum {
    double s, c;
};

__device__ __forceinline__ void csum_add(CSum& a, double v) {
    const double t = a.s + v;
    const double bp = t - a.s;
    const double e = (a.s - (t - bp)) + (v - bp);
    a.s = t;
    a.c += e;
}

__device__ __forceinline__ CSum csum_merge(CSum a, CSum b) {
    const double t = a.s + b.s;
    const double bp = t - a.s;
    const double e = (a.s - (t - bp)) + (b.s - bp);
    CSum r;
    r.s = t;
    r.c = a.c + b.c + e;
    return r;
}

// K compensated sums per thread.
template <int K>
struct Acc {
    CSum v[K];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int k = 0; k < K; ++k) v[k].s = v[k].c = 0.0;
    }
};

template <int K>
__device__ __forceinline__ void acc_merge_into(Acc<K>& a, const Acc<K>& b) {
#pragma unroll
    for (int k = 0; k < K; ++k) a.v[k] = csum_merge(a.v[k], b.v[k]);
}

template <int K>
__device__ __forceinline__ Acc<K> wave_reduce(Acc<K> a) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        Acc<K> o;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            o.v[k].s = __shfl_xor(a.v[k].s, off, kWave);
            o.v[k].c = __shfl_xor(a.v[k].c, off, kWave);
        }
        acc_merge_into(a, o);
    }
    return a;
}

// Block reduction; the result is valid in thread 0.  `lds` holds
// (blockDim/64) * K CSums.  Fixed order => deterministic for a launch shape.
template <int K>
__device__ __forceinline__ Acc<K> block_reduce(Acc<K> a, CSum* lds) {
    a = wave_reduce<K>(a);
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int nwaves = (blockDim.x + kWave - 1) / kWave;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) lds[wave * K + k] = a.v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < nwaves; ++w) {
#pragma unroll
            for (int k = 0; k < K; ++k) a.v[k] = csum_merge(a.v[k], lds[w * K + k]);
        }
    }
    return a;
}

// Cross-block data (partials, counters, flags) is published with coherent
// agent-scope atomic stores, which write through the per-XCD L2, and read
// with coherent atomic loads.  Ordering a block's published values before its
// arrival only needs those stores to be complete (vmcnt 0), so no agent-scope
// fence -- i.e. no whole-L2 write-back or invalidate -- is needed anywhere.
__device__ __forceinline__ void publish(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double fetch(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void publish_u32(unsigned int* p, unsigned int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned int fetch_u32(const unsigned int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long fetch_u64(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Wait until every memory operation this wave issued is complete.
__device__ __forceinline__ void stores_complete() { __builtin_amdgcn_s_waitcnt(0); }

// Arrival counting of grid_reduce.  Device-scope atomics on one address run
// one after another, so a single arrival counter costs nblk such atomics in
// sequence at the end of a launch whose blocks finish together.  Block b
// counts on group counter b mod kArriveGroups instead; the last block of each
// group then counts on a top counter, whose last arrival is the last block
// overall.  Each counter has its own 128-byte line; one reduction uses
// kCounterWords uint32 of counter space.
constexpr int kArriveGroups = 32;
constexpr int kCounterLine = 32;  // uint32 per 128-byte line
constexpr int kCounterWords = (kArriveGroups + 1) * kCounterLine;

// Per-block partial write + "last block finishes" grid reduction.
//   partials: nblk * K CSums;  counter: kCounterWords uint32, zero at launch and left
//   zero on exit.  Returns true in the (single) block that performed the
//   final pass, whose thread 0 then holds the totals in `tot`.
//   blk / nblk: this block's index and the block count of the reduction (the
//   launch grid for a single request; the request's share of a batched grid).
template <int K>
__device__ __forceinline__ bool grid_reduce(Acc<K> a, CSum* __restrict__ partials,
                                            unsigned int* __restrict__ counter, CSum* lds,
                                            Acc<K>& tot, unsigned int blk, unsigned int nblk) {
    __shared__ bool is_last;
    a = block_reduce<K>(a, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            publish(&partials[(size_t)blk * K + k].s, a.v[k].s);
            publish(&partials[(size_t)blk * K + k].c, a.v[k].c);
        }
        stores_complete();
        // Group g holds blocks g, g + 32, ...: ceil((nblk - g) / 32) of them
        // (blk < nblk, so g < nblk and the group is not empty).
        const unsigned int g = blk % kArriveGroups;
        const unsigned int ng = (nblk - g + kArriveGroups - 1) / kArriveGroups;
        const unsigned int groups = nblk < (unsigned)kArriveGroups ? nblk : (unsigned)kArriveGroups;
        bool last = false;
        if (atomicAdd(&counter[g * kCounterLine], 1u) == ng - 1)
            last = atomicAdd(&counter[kArriveGroups * kCounterLine], 1u) == groups - 1;
        is_last = last;
    }
    __syncthreads();
    if (!is_last) return false;
    // Thread t folds partials t, t+B, t+2B, ... in order, then a block reduce.
    // The loads of U consecutive partials are issued before their merges, so
    // they are in flight together; the merge order is unchanged.
    constexpr int U = K <= 2 ? 8 : K <= 4 ? 4 : K <= 8 ? 2 : 1;
    Acc<K> r;
    r.zero();
    unsigned int b = threadIdx.x;
    for (; b + (U - 1) * blockDim.x < nblk; b += U * blockDim.x) {
        CSum p[U][K];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < K; ++k) {
                p[u][k].s = fetch(&partials[(size_t)(b + u * blockDim.x) * K + k].s);
                p[u][k].c = fetch(&partials[(size_t)(b + u * blockDim.x) * K + k].c);
            }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < K; ++k) r.v[k] = csum_merge(r.v[k], p[u][k]);
    }
    for (; b < nblk; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            CSum p;
            p.s = fetch(&partials[(size_t)b * K + k].s);
            p.c = fetch(&partials[(size_t)b * K + k].c);
            r.v[k] = csum_merge(r.v[k], p);
        }
    }
    __syncthreads();  // lds reuse
    r = block_reduce<K>(r, lds);
    // Every block has counted (the top counter is complete): clear the
    // counters for the next launch.
    for (unsigned int c = threadIdx.x; c <= (unsigned)kArriveGroups; c += blockDim.x)
        publish_u32(&counter[c * kCounterLine], 0u);
    if (threadIdx.x == 0) tot = r;
    return true;
}

// Final FP64 moments of NP (sum, sum of squares) pairs: out[4k..4k+3] =
// sum, sum of squares, mean, sample variance.
template <int NP>
__device__ __forceinline__ void write_moments(const Acc<2 * NP>& tot, int64_t N, double* out) {
    const double n = (double)N;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const double sum = tot.v[2 * k].s + tot.v[2 * k].c;
        const double sq = tot.v[2 * k + 1].s + tot.v[2 * k + 1].c;
        const double mean = sum / n;
        out[4 * k + 0] = sum;
        out[4 * k + 1] = sq;
        out[4 * k + 2] = mean;
        out[4 * k + 3] = N > 1 ? (sq - sum * mean) / (n - 1.0) : 0.0;
    }
}

// Batched launches: job j owns blocks [first_block, first_block + nblk) of the
// grid.  Returns the job of block b (jobs sorted by first_block; b is
// block-uniform, so the search runs on scalar loads).
template <typename Job>
__device__ __forceinline__ int find_job(const Job* __restrict__ jobs, int njobs, unsigned int b) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first_block <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace pcf
