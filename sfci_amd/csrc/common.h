// Device helpers shared by the SFCI HIP kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SFCI_WAVE 64
#define SFCI_BLOCK 256
#define SFCI_MAX_ORB 32          // spatial orbitals with 64-bit determinant keys
#define SFCI_MAX_K 32            // electrons per spin

// ---------------------------------------------------------------------------------
// Exact 128-bit fixed-point accumulation, unit 2^-84.
//
// Terms are cut to the 2^-84 grid by truncating the magnitude toward zero; sums are
// exact two's-complement integer sums (order-independent); the final value is the
// sum rounded to nearest-even double.  |term| must stay below 2^42 (checked by the
// callers through fx_in_range).
// ---------------------------------------------------------------------------------
struct fx128 {
    uint64_t lo;
    uint64_t hi;   // two's complement high word
};

__host__ __device__ __forceinline__ fx128 fx_zero() { return fx128{0, 0}; }

__host__ __device__ __forceinline__ fx128 fx_add(fx128 a, fx128 b) {
    fx128 r;
    r.lo = a.lo + b.lo;
    r.hi = a.hi + b.hi + (r.lo < a.lo ? 1ull : 0ull);
    return r;
}

__host__ __device__ __forceinline__ fx128 fx_neg(fx128 a) {
    fx128 r;
    r.lo = ~a.lo + 1ull;
    r.hi = ~a.hi + (r.lo == 0 ? 1ull : 0ull);
    return r;
}

__host__ __device__ __forceinline__ bool fx_in_range(double x) {
    return fabs(x) < 4398046511104.0;   // 2^42
}

__host__ __device__ __forceinline__ fx128 fx_from(double x) {
    uint64_t b;
    memcpy(&b, &x, 8);
    uint32_t e = (uint32_t)(b >> 52) & 0x7ffu;
    if (e == 0 || e == 0x7ffu) return fx_zero();          // zero, subnormal (< 2^-84), inf/nan
    uint64_t M = (b & 0xfffffffffffffull) | (1ull << 52);
    int s = (int)e - 991;                                   // |x| * 2^84 = M * 2^s
    fx128 r;
    if (s >= 64) {
        r.lo = 0;
        r.hi = M << (s - 64);
    } else if (s > 0) {
        r.lo = M << s;
        r.hi = M >> (64 - s);
    } else if (s == 0) {
        r.lo = M;
        r.hi = 0;
    } else if (s > -64) {
        r.lo = M >> (-s);
        r.hi = 0;
    } else {
        r.lo = 0;
        r.hi = 0;
    }
    return (b >> 63) ? fx_neg(r) : r;
}

// Limb form of the same fixed-point value for short exact sums: a term x with |x| < 2^42
// contributes l2 = trunc(x), l1 = trunc(frac(x) * 2^42), l0 = trunc(rest * 2^42) (exact
// FP64 steps), each with the sign of x and |l1|, |l0| < 2^42, so that l2 2^84 + l1 2^42 +
// l0 is the integer fx_from(x) represents.  Up to 2^21 terms sum without overflow;
// fxl_join gives the fx128 value.
struct FxLimbs {
    long long l0, l1, l2;
};

__host__ __device__ __forceinline__ FxLimbs fxl_zero() { return FxLimbs{0, 0, 0}; }

// adds x; returns false (adding nothing) when |x| >= 2^42 or x is not finite
__host__ __device__ __forceinline__ bool fxl_add(FxLimbs& a, double x) {
    if (!(fabs(x) < 4398046511104.0)) return false;
    const double t2 = trunc(x);
    const double r1 = (x - t2) * 0x1p42;
    const double t1 = trunc(r1);
    const double t0 = trunc((r1 - t1) * 0x1p42);
    a.l2 += (long long)t2;
    a.l1 += (long long)t1;
    a.l0 += (long long)t0;
    return true;
}

__host__ __device__ __forceinline__ fx128 fxl_join(const FxLimbs& a) {
    const uint64_t l0 = (uint64_t)a.l0, l1 = (uint64_t)a.l1, l2 = (uint64_t)a.l2;
    fx128 t{l0, a.l0 < 0 ? ~0ull : 0ull};
    t = fx_add(t, fx128{l1 << 42, (uint64_t)(a.l1 >> 22)});
    return fx_add(t, fx128{0, l2 << 20});
}

__host__ __device__ __forceinline__ int fx_clz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __clzll((long long)x);
#else
    return __builtin_clzll(x);
#endif
}

__host__ __device__ __forceinline__ double fx_bits_to_double(uint64_t b) {
    double d;
    memcpy(&d, &b, 8);
    return d;
}

// Round-to-nearest-even conversion of the exact sum.
__host__ __device__ __forceinline__ double fx_to(fx128 v) {
    bool neg = (int64_t)v.hi < 0;
    if (neg) v = fx_neg(v);
    if (v.hi == 0 && v.lo == 0) return 0.0;
    int p = v.hi ? 127 - fx_clz64(v.hi) : 63 - fx_clz64(v.lo);
    double r;
    if (p <= 52) {
        r = (double)v.lo * 0x1p-84;                             // exact
    } else {
        int sh = p - 52;
        uint64_t m, rem_hi, rem_lo, half_hi, half_lo;
        // m = v >> sh
        if (sh >= 64) m = v.hi >> (sh - 64);
        else m = (v.lo >> sh) | (v.hi << (64 - sh));
        // rem = v & (2^sh - 1); half = 2^(sh-1)
        if (sh >= 64) {
            rem_lo = v.lo;
            rem_hi = (sh == 64) ? 0 : (v.hi & ((1ull << (sh - 64)) - 1));
        } else {
            rem_lo = v.lo & ((1ull << sh) - 1);
            rem_hi = 0;
        }
        if (sh - 1 >= 64) { half_hi = 1ull << (sh - 1 - 64); half_lo = 0; }
        else { half_hi = 0; half_lo = 1ull << (sh - 1); }
        bool gt = rem_hi > half_hi || (rem_hi == half_hi && rem_lo > half_lo);
        bool eq = rem_hi == half_hi && rem_lo == half_lo;
        if (gt || (eq && (m & 1ull))) ++m;
        if (m == (1ull << 53)) { m >>= 1; ++sh; }
        // value = m * 2^(sh - 84), m in [2^52, 2^53)
        uint64_t expo = (uint64_t)(sh - 84 + 52 + 1023);
        r = fx_bits_to_double((expo << 52) | (m & 0xfffffffffffffull));
    }
    return neg ? -r : r;
}

#if defined(__HIPCC__)
// Order-independent device-wide accumulation of a 128-bit value.
__device__ __forceinline__ void fx_atomic_add(unsigned long long* acc, fx128 v) {
    unsigned long long old = atomicAdd(acc, (unsigned long long)v.lo);
    unsigned long long carry = (old + v.lo) < old ? 1ull : 0ull;
    atomicAdd(acc + 1, (unsigned long long)(v.hi + carry));
}

// Block-wide reduction of one fx128 per thread, then one atomic pair per block.
template <int BLOCK>
__device__ __forceinline__ void fx_block_reduce_atomic(fx128 v, unsigned long long* acc) {
    __shared__ uint64_t s_lo[BLOCK / SFCI_WAVE], s_hi[BLOCK / SFCI_WAVE];
    // wave reduction via shuffles
    for (int off = SFCI_WAVE / 2; off > 0; off >>= 1) {
        fx128 o;
        o.lo = __shfl_down(v.lo, off, SFCI_WAVE);
        o.hi = __shfl_down(v.hi, off, SFCI_WAVE);
        v = fx_add(v, o);
    }
    int lane = threadIdx.x & (SFCI_WAVE - 1), w = threadIdx.x / SFCI_WAVE;
    if (lane == 0) { s_lo[w] = v.lo; s_hi[w] = v.hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        fx128 t = fx_zero();
        for (int i = 0; i < BLOCK / SFCI_WAVE; ++i) t = fx_add(t, fx128{s_lo[i], s_hi[i]});
        if (t.lo | t.hi) fx_atomic_add(acc, t);
    }
    __syncthreads();
}
#endif

// ---------------------------------------------------------------------------------
// Determinant bit strings.  Bit p is spin orbital p; even bits are spin 0 (the spin
// the reference ranks, src/bitstr.cpp:80), odd bits spin 1.  Spatial orbital i of spin
// s is bit 2*i + s.
// ---------------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint32_t spin_string(uint64_t d, int s) {
    // compress the bits 2*i + s into bit i
    uint64_t x = (d >> s) & 0x5555555555555555ull;
    x = (x | (x >> 1)) & 0x3333333333333333ull;
    x = (x | (x >> 2)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x >> 4)) & 0x00ff00ff00ff00ffull;
    x = (x | (x >> 8)) & 0x0000ffff0000ffffull;
    x = (x | (x >> 16)) & 0x00000000ffffffffull;
    return (uint32_t)x;
}

__host__ __device__ __forceinline__ uint64_t spread_string(uint32_t s32) {
    uint64_t x = s32;
    x = (x | (x << 16)) & 0x0000ffff0000ffffull;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

__host__ __device__ __forceinline__ int popc64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __popcll(x);
#else
    return __builtin_popcountll(x);
#endif
}

__host__ __device__ __forceinline__ int ctz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __ffsll((long long)x) - 1;
#else
    return __builtin_ctzll(x);
#endif
}

// (-1)^(number of occupied spin orbitals strictly between p and q)
__host__ __device__ __forceinline__ double sign_between(uint64_t d, int p, int q) {
    int lo = p < q ? p : q, hi = p < q ? q : p;
    uint64_t mask = (hi - lo > 1) ? ((((uint64_t)1 << (hi - lo - 1)) - 1) << (lo + 1)) : 0ull;
    return (popc64(d & mask) & 1) ? -1.0 : 1.0;
}

// Multiplicative hash of a target index; the top hb bits select a bucket, the bits
// below them a slot inside the bucket's table.
__host__ __device__ __forceinline__ uint64_t target_hash_full(uint64_t t) { return t * 0x9E3779B97F4A7C15ull; }
__host__ __device__ __forceinline__ uint64_t target_hash(uint64_t t, int hb) {
    return hb ? (target_hash_full(t) >> (64 - hb)) : 0ull;
}

// colex rank of a k-subset given as a bit string over spatial orbitals:
// sum_t C(pos_t, t+1), pos_t ascending.  binom is a row-major table with 33 columns (the
// host's [65][33] in 64 bits, the device's [33][33] in 32 bits).
template <typename BT>
__host__ __device__ __forceinline__ uint64_t colex_rank(uint32_t s, const BT* binom) {
    uint64_t r = 0;
    int t = 1;
    while (s) {
        int pos = ctz64(s);
        r += binom[pos * 33 + t];
        s &= s - 1;
        ++t;
    }
    return r;
}

// inverse of colex_rank for k-subsets of n_orb orbitals
template <typename BT>
__host__ __device__ __forceinline__ uint32_t colex_unrank(uint64_t r, int k, int n_orb, const BT* binom) {
    uint32_t s = 0;
    int pos = n_orb - 1;
    for (int t = k; t >= 1; --t) {
        while (pos > 0 && binom[pos * 33 + t] > r) --pos;
        s |= 1u << pos;
        r -= binom[pos * 33 + t];
        --pos;
    }
    return s;
}
