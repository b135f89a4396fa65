// H.v on 128-bit determinant keys: more than 64 spin orbitals, up to 128.
//
// No reference counterpart: the reference keeps a determinant in one 64-bit word
// (dynVec<u64,double>, src/all.h:150-194; num_orb <= 64, src/main.cpp:84-85).  The
// semantics are those of SpMVMtply_bucket (src/functions.cpp:511-804) as sfci_hv
// implements them: top-Nd input by |c|, y = coeff (H - lambda) kept with every sum exact,
// diag = <kept|y> before the cut, strict |y| > cut, top out_cap by (|y| desc, key asc),
// ascending key order.  A key is two words, lo = spin orbitals 0-63 and hi = spin
// orbitals 64-127, and keys compare as the 128-bit integer hi:lo.
//
// Pipeline of one call:
//   1. input  : top-Nd by |c| (stable sort of the key-ordered input by |c|, the first nd
//               flagged and compacted back in key order).
//   2. passes : targets are split into n_pass classes by a 64-bit hash of the key, so
//               that one class's candidates fit the workspace.
//   3. per pass
//      generate one workgroup per source: the diagonal, singles, same-spin and
//               opposite-spin doubles, each connection evaluated once with its full
//               Slater-Condon element (exact limb sums); a connection is kept when its
//               hash class is the pass, appended as (lo, hi, value) with one atomic per
//               wavefront
//      group    stable LSD radix sort on a 48-bit hash of the key (index payload), so
//               that the records of one key are adjacent (keys that share a hash form
//               one run and are told apart inside it)
//      reduce   exact fixed-point sum per distinct key; diag from the sources whose own
//               key falls in the pass; |y| > cut appended to the staging list
//   4. output : staging list sorted by the 128-bit key (LSD on lo, then on hi), top
//               out_cap.
#include "internal.h"

namespace sfci {

namespace {

constexpr int W_BLOCK = 256;

__host__ __device__ __forceinline__ uint64_t wide_hash(uint64_t lo, uint64_t hi) {
    uint64_t z = lo * 0x9E3779B97F4A7C15ull ^ (hi + 0x632BE59BD9B4E019ull) * 0xC2B2AE3D27D4EB4Full;
    z ^= z >> 31;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 29;
    return z;
}

// hash class of a key among n_pass classes
__host__ __device__ __forceinline__ uint32_t wide_class(uint64_t lo, uint64_t hi, uint32_t n_pass) {
    return (uint32_t)(((wide_hash(lo, hi) >> 32) * (uint64_t)n_pass) >> 32);
}

__device__ __forceinline__ bool bit_of(uint64_t lo, uint64_t hi, int p) {
    return p < 64 ? ((lo >> p) & 1ull) : ((hi >> (p - 64)) & 1ull);
}

__device__ __forceinline__ void flip(uint64_t& lo, uint64_t& hi, int p) {
    if (p < 64) lo ^= 1ull << p;
    else hi ^= 1ull << (p - 64);
}

// mask of bits [l, h) inside the 64-bit word that starts at bit w0 (0 <= l <= h <= 128)
__device__ __forceinline__ uint64_t word_mask(int l, int h, int w0) {
    const int a = max(l, w0), b = min(h, w0 + 64);
    if (a >= b) return 0ull;
    const int w = b - a;
    return (w == 64 ? ~0ull : ((1ull << w) - 1ull)) << (a - w0);
}

// (-1)^(number of occupied spin orbitals strictly between p and q)
__device__ __forceinline__ double sign_between_w(uint64_t lo, uint64_t hi, int p, int q) {
    const int a = min(p, q) + 1, b = max(p, q);
    const int c = __popcll(lo & word_mask(a, b, 0)) + __popcll(hi & word_mask(a, b, 64));
    return (c & 1) ? -1.0 : 1.0;
}

__device__ __forceinline__ double g_w(const HamDev& H, int m, int n, int p, int q) {
    if (((m ^ n) & 1) | ((p ^ q) & 1)) return 0.0;
    const size_t N = (size_t)H.n_orb;
    const size_t blk = (size_t)(((m & 1) << 1) | (p & 1));
    return H.g[blk * H.g_stride + (((size_t)(m >> 1) * N + (n >> 1)) * N + (p >> 1)) * N + (q >> 1)];
}

__device__ __forceinline__ double h_w(const HamDev& H, int m, int n) {
    if ((m ^ n) & 1) return 0.0;
    const size_t N = (size_t)H.n_orb;
    return H.h[(size_t)(m & 1) * N * N + (size_t)(m >> 1) * N + (n >> 1)];
}

struct WSrc {
    uint64_t lo, hi;
    double temp;        // coeff * c_j
    int n_occ_so;
    int ok;             // the determinant has k electrons of each spin inside n_so
    uint8_t occ_so[64]; // occupied spin orbitals, ascending
    uint8_t occ[2][32]; // occupied spatial orbitals of each spin, ascending
    uint8_t vir[2][64]; // empty spatial orbitals of each spin, ascending
};

// <D''|H|D>, D'' = D - {n,q} + {x,y}, n < q occupied, x < y empty (spin orbitals)
__device__ __forceinline__ double double_elem_w(const HamDev& H, uint64_t lo, uint64_t hi, int n, int q, int x,
                                                int y) {
    uint64_t l1 = lo, h1 = hi;
    flip(l1, h1, q);
    flip(l1, h1, y);
    const double s1 = sign_between_w(lo, hi, q, y) * sign_between_w(l1, h1, n, x);
    uint64_t l2 = lo, h2 = hi;
    flip(l2, h2, q);
    flip(l2, h2, x);
    const double s2 = sign_between_w(lo, hi, q, x) * sign_between_w(l2, h2, n, y);
    return s1 * g_w(H, x, n, y, q) + s2 * g_w(H, y, n, x, q);
}

template <bool MIXED>
__device__ __forceinline__ double wide_val(double hij, double temp) {
    if (MIXED) return (double)((float)hij * (float)temp);
    return hij * temp;
}

// Where a source's records go.  Single pass: record `idx` of source j at j * n_conn + idx
// (every connection is kept; no atomics).  Several passes: the source's in-pass records
// are counted first and reserved with one global atomic (base); each wavefront then takes
// its slots from a workgroup cursor in LDS with one LDS atomic.
struct WEmit {
    uint64_t* rlo;
    uint64_t* rhi;
    double* rv;
    uint64_t cap;
    uint64_t base;
    unsigned* cursor;   // LDS; nullptr in single-pass mode
    uint32_t* flags;
    // called by every lane of a wavefront that is still in the enclosing loop
    __device__ __forceinline__ void put(bool want, uint32_t idx, uint64_t lo, uint64_t hi, double val) const {
        uint64_t i;
        if (!cursor) {
            if (!want) return;
            i = base + idx;
        } else {
            const uint64_t bal = __ballot(want);
            if (bal == 0) return;
            const int lane = threadIdx.x & (SFCI_WAVE - 1);
            const int leader = __ffsll((long long)bal) - 1;
            unsigned off = 0;
            if (lane == leader) off = atomicAdd(cursor, (unsigned)__popcll(bal));
            off = __shfl(off, leader, SFCI_WAVE);
            if (!want) return;
            const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
            i = base + off + (uint64_t)__popcll(bal & lt);
        }
        if (i < cap) {
            rlo[i] = lo;
            rhi[i] = hi;
            rv[i] = val;
        } else {
            atomicOr(flags, FLAG_GEN_OVERFLOW);
        }
    }
};

// Target of connection e of the staged source (e >= 1; 0 is the diagonal), in the order
// singles (spin 0, spin 1), same-spin doubles (spin 0, spin 1), opposite-spin doubles.
__device__ __forceinline__ void conn_target(const HamDev& H, const WSrc& s, uint32_t e, uint64_t& tl,
                                            uint64_t& th) {
    const uint32_t kv = (uint32_t)H.n_sa, dss = (uint32_t)H.n_daa, v = (uint32_t)H.v;
    tl = s.lo;
    th = s.hi;
    uint32_t x = e - 1;
    if (x < 2 * kv) {
        const int sp = x >= kv ? 1 : 0;
        x -= (uint32_t)sp * kv;
        const uint32_t i = x / v, a = x - i * v;
        flip(tl, th, 2 * s.occ[sp][i] + sp);
        flip(tl, th, 2 * s.vir[sp][a] + sp);
        return;
    }
    x -= 2 * kv;
    if (x < 2 * dss) {
        const int sp = x >= dss ? 1 : 0;
        x -= (uint32_t)sp * dss;
        const uint32_t op = x / H.n_vpair, vp = x - op * H.n_vpair;
        const uint16_t po = H.occ_pairs[op], pv = H.vir_pairs[vp];
        flip(tl, th, 2 * s.occ[sp][po & 0xff] + sp);
        flip(tl, th, 2 * s.occ[sp][po >> 8] + sp);
        flip(tl, th, 2 * s.vir[sp][pv & 0xff] + sp);
        flip(tl, th, 2 * s.vir[sp][pv >> 8] + sp);
        return;
    }
    x -= 2 * dss;
    const uint32_t s0 = x / kv, s1 = x - s0 * kv;
    const uint32_t i = s0 / v, a = s0 - i * v, jj = s1 / v, b = s1 - jj * v;
    flip(tl, th, 2 * s.occ[0][i]);
    flip(tl, th, 2 * s.vir[0][a]);
    flip(tl, th, 2 * s.occ[1][jj] + 1);
    flip(tl, th, 2 * s.vir[1][b] + 1);
}

// Connections of each source whose hash class is `pass`: (key, coeff (H - lambda)_tj c_j)
// for the diagonal, H_tj coeff c_j for every other connected t.
template <bool MIXED>
__global__ __launch_bounds__(W_BLOCK) void k_wide_generate(HamDev H, int n_so, const uint64_t* __restrict__ slo,
                                                           const uint64_t* __restrict__ shi,
                                                           const double* __restrict__ sv, size_t ns, double lambda,
                                                           double coeff, uint32_t n_pass, uint32_t pass,
                                                           uint64_t* __restrict__ rlo, uint64_t* __restrict__ rhi,
                                                           double* __restrict__ rv, uint64_t cap,
                                                           unsigned long long* count, uint32_t* flags) {
    __shared__ WSrc s;
    __shared__ unsigned s_cnt, s_cur;
    __shared__ unsigned long long s_base;
    const int tid = threadIdx.x, lane = tid & (SFCI_WAVE - 1), wv = tid / SFCI_WAVE;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int N = H.n_orb, k = H.k, v = H.v;
    const uint32_t kv = (uint32_t)H.n_sa;
    for (size_t j = blockIdx.x; j < ns; j += gridDim.x) {
        __syncthreads();   // the previous source's lists are no longer read
        if (wv == 0) {
            const uint64_t lo = slo[j], hi = shi[j];
            int n_occ = 0;
            bool ok = true;
            for (int half = 0; half < 2; ++half) {
                const int p = half * 64 + lane;
                const bool oc = bit_of(lo, hi, p);
                if (oc && p >= n_so) ok = false;
                const uint64_t b = __ballot(oc && p < n_so);
                const int t = n_occ + __popcll(b & lt);
                if (oc && p < n_so && t < 64) s.occ_so[t] = (uint8_t)p;
                n_occ += __popcll(b);
            }
            for (int sp = 0; sp < 2; ++sp) {
                const bool in = lane < N;
                const bool oc = in && bit_of(lo, hi, 2 * lane + sp);
                const uint64_t b_oc = __ballot(oc), b_vi = __ballot(in && !oc);
                if ((int)__popcll(b_oc) != k) ok = false;
                const int t = __popcll(b_oc & lt);
                if (oc && t < 32) s.occ[sp][t] = (uint8_t)lane;
                else if (in && !oc) s.vir[sp][__popcll(b_vi & lt)] = (uint8_t)lane;
            }
            if (!__all(ok)) ok = false;
            if (lane == 0) {
                s.lo = lo;
                s.hi = hi;
                s.temp = coeff * sv[j];
                s.n_occ_so = n_occ;
                s.ok = ok ? 1 : 0;
                if (!ok) atomicOr(flags, FLAG_BAD_INPUT);
            }
        }
        __syncthreads();
        if (!s.ok) continue;   // block-uniform
        const uint64_t lo = s.lo, hi = s.hi;
        const double temp = s.temp;
        const int n_occ = s.n_occ_so;
        const uint32_t n_conn = (uint32_t)H.n_conn;
        bool bad = false;
        WEmit em{rlo, rhi, rv, cap, (uint64_t)j * n_conn, nullptr, flags};
        if (n_pass > 1) {
            // count this source's in-pass connections, reserve them with one global atomic
            unsigned cnt = 0;
            for (uint32_t e = tid; e < n_conn; e += W_BLOCK) {
                uint64_t tl = lo, th = hi;
                if (e) conn_target(H, s, e, tl, th);
                cnt += wide_class(tl, th, n_pass) == pass ? 1u : 0u;
            }
            for (int off = SFCI_WAVE / 2; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, SFCI_WAVE);
            if (tid == 0) {
                s_cnt = 0;
                s_cur = 0;
            }
            __syncthreads();
            if (lane == 0) atomicAdd(&s_cnt, cnt);
            __syncthreads();
            if (tid == 0) s_base = atomicAdd(count, (unsigned long long)s_cnt);
            __syncthreads();
            em.base = s_base;
            em.cursor = &s_cur;
        }
        // diagonal: E_core + sum_occ h_mm + sum_{m != p} 1/2 [(mm|pp) - (mp|pm)], exact
        if (wv == 0) {
            const bool mine = wide_class(lo, hi, n_pass) == pass;   // wave-uniform
            if (mine) {
                FxLimbs acc = fxl_zero();
                bool dok = true;
                for (int t = lane; t < n_occ * n_occ; t += SFCI_WAVE) {
                    const int a = t / n_occ, b = t - a * n_occ;
                    const int m = s.occ_so[a];
                    if (a == b) {
                        dok &= fxl_add(acc, h_w(H, m, m));
                    } else {
                        const int p = s.occ_so[b];
                        dok &= fxl_add(acc, 0.5 * g_w(H, m, m, p, p));
                        dok &= fxl_add(acc, -(0.5 * g_w(H, m, p, p, m)));
                    }
                }
                fx128 f = fxl_join(acc);
                for (int off = SFCI_WAVE / 2; off > 0; off >>= 1) {
                    fx128 o;
                    o.lo = __shfl_down(f.lo, off, SFCI_WAVE);
                    o.hi = __shfl_down(f.hi, off, SFCI_WAVE);
                    f = fx_add(f, o);
                }
                dok = __all(dok);
                double dv = 0.0;
                if (lane == 0) {
                    const double e = fx_to(fx_add(fx_from(H.constant), f));
                    dv = (e - lambda) * temp;
                    if (!dok || !fx_in_range(dv)) bad = true;
                }
                em.put(lane == 0, 0, lo, hi, dv);
            }
        }
        // singles, both spins: spin orbital n -> m
        for (uint32_t e = tid; e < 2 * kv; e += W_BLOCK) {
            const int sp = e >= kv ? 1 : 0;
            const uint32_t x = e - (uint32_t)sp * kv;
            const uint32_t i = x / (uint32_t)v, a = x - i * (uint32_t)v;
            const int n = 2 * s.occ[sp][i] + sp, m = 2 * s.vir[sp][a] + sp;
            uint64_t tl = lo, th = hi;
            flip(tl, th, n);
            flip(tl, th, m);
            const bool want = wide_class(tl, th, n_pass) == pass;
            double val = 0.0;
            if (want) {
                FxLimbs acc = fxl_zero();
                bool ok = fxl_add(acc, h_w(H, m, n));
                for (int b = 0; b < n_occ; ++b) {
                    const int p = s.occ_so[b];
                    if (p == n) continue;
                    ok &= fxl_add(acc, g_w(H, m, n, p, p));
                    ok &= fxl_add(acc, -g_w(H, m, p, p, n));
                }
                const double hij = ok ? sign_between_w(lo, hi, m, n) * fx_to(fxl_join(acc)) : __builtin_nan("");
                val = wide_val<MIXED>(hij, temp);
                bad |= !fx_in_range(val);
            }
            em.put(want, 1 + e, tl, th, val);
        }
        // same-spin doubles, both spins: occupied pair (i < j) -> empty pair (a < b)
        const uint32_t dss = (uint32_t)H.n_daa;
        for (uint32_t e = tid; e < 2 * dss; e += W_BLOCK) {
            const int sp = e >= dss ? 1 : 0;
            const uint32_t x = e - (uint32_t)sp * dss;
            const uint32_t op = x / H.n_vpair, vp = x - op * H.n_vpair;
            const uint16_t po = H.occ_pairs[op], pv = H.vir_pairs[vp];
            const int n = 2 * s.occ[sp][po & 0xff] + sp, q = 2 * s.occ[sp][po >> 8] + sp;
            const int xx = 2 * s.vir[sp][pv & 0xff] + sp, y = 2 * s.vir[sp][pv >> 8] + sp;
            uint64_t tl = lo, th = hi;
            flip(tl, th, n);
            flip(tl, th, q);
            flip(tl, th, xx);
            flip(tl, th, y);
            const bool want = wide_class(tl, th, n_pass) == pass;
            double val = 0.0;
            if (want) {
                val = wide_val<MIXED>(double_elem_w(H, lo, hi, n, q, xx, y), temp);
                bad |= !fx_in_range(val);
            }
            em.put(want, 1 + 2 * kv + e, tl, th, val);
        }
        // opposite-spin doubles: spin-0 i -> a with spin-1 j -> b
        const uint32_t dos = kv * kv;
        for (uint32_t e = tid; e < dos; e += W_BLOCK) {
            const uint32_t s0 = e / kv, s1 = e - s0 * kv;
            const uint32_t i = s0 / (uint32_t)v, a = s0 - i * (uint32_t)v;
            const uint32_t jj = s1 / (uint32_t)v, b = s1 - jj * (uint32_t)v;
            const int ni = 2 * s.occ[0][i], xa = 2 * s.vir[0][a];
            const int nj = 2 * s.occ[1][jj] + 1, yb = 2 * s.vir[1][b] + 1;
            const int n = min(ni, nj), q = max(ni, nj), xx = min(xa, yb), y = max(xa, yb);
            uint64_t tl = lo, th = hi;
            flip(tl, th, n);
            flip(tl, th, q);
            flip(tl, th, xx);
            flip(tl, th, y);
            const bool want = wide_class(tl, th, n_pass) == pass;
            double val = 0.0;
            if (want) {
                val = wide_val<MIXED>(double_elem_w(H, lo, hi, n, q, xx, y), temp);
                bad |= !fx_in_range(val);
            }
            em.put(want, 1 + 2 * kv + 2 * dss + e, tl, th, val);
        }
        if (bad) atomicOr(flags, FLAG_RANGE);
    }
}

// ---- small elementwise kernels ----------------------------------------------------
__global__ void k_wide_split(const uint64_t* __restrict__ k2, const double* __restrict__ v, size_t n,
                             uint64_t* __restrict__ lo, uint64_t* __restrict__ hi, double* __restrict__ vo) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        lo[i] = k2[2 * i];
        hi[i] = k2[2 * i + 1];
        vo[i] = v[i];
    }
}

__global__ void k_wide_join(const uint64_t* __restrict__ lo, const uint64_t* __restrict__ hi,
                            const double* __restrict__ v, size_t n, uint64_t* __restrict__ k2,
                            double* __restrict__ vo) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        k2[2 * i] = lo[i];
        k2[2 * i + 1] = hi[i];
        vo[i] = v[i];
    }
}

__global__ void k_wide_iota(uint64_t* __restrict__ x, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] = i;
}

template <typename T>
__global__ void k_wide_gather(const T* __restrict__ src, const uint64_t* __restrict__ idx, size_t n,
                              T* __restrict__ dst) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

// ranking key: ascending order = |v| descending (bit pattern of |v| inverted)
__global__ void k_wide_rank_key(const double* __restrict__ v, size_t n, uint64_t* __restrict__ key) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint64_t b;
        const double x = v[i];
        __builtin_memcpy(&b, &x, 8);
        key[i] = 0x7fffffffffffffffull - (b & 0x7fffffffffffffffull);
    }
}

__global__ void k_wide_mark(const uint64_t* __restrict__ idx, size_t cap, uint32_t* __restrict__ fl) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (size_t)gridDim.x * blockDim.x)
        fl[idx[i]] = 1u;
}

// start[pos[i]] = i for every head; start[n_seg] = n
__global__ void k_wide_seg_start(const uint32_t* __restrict__ head, const uint64_t* __restrict__ pos, size_t n,
                                 uint64_t* __restrict__ start) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (head[i]) start[pos[i]] = i;
        if (i == n - 1) start[pos[i] + head[i]] = n;
    }
}

__device__ __forceinline__ bool key_less(uint64_t alo, uint64_t ahi, uint64_t blo, uint64_t bhi) {
    return ahi < bhi || (ahi == bhi && alo < blo);
}

__global__ void k_wide_flag_above(const double* __restrict__ y, size_t n, double cut, uint32_t* __restrict__ fl) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        fl[i] = fabs(y[i]) > cut ? 1u : 0u;
}

__global__ void k_wide_compact(const uint32_t* __restrict__ fl, const uint64_t* __restrict__ pos, size_t n,
                               const uint64_t* __restrict__ lo, const uint64_t* __restrict__ hi,
                               const double* __restrict__ y, uint64_t* __restrict__ dlo, uint64_t* __restrict__ dhi,
                               double* __restrict__ dy) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (!fl[i]) continue;
        const uint64_t p = pos[i];
        dlo[p] = lo[i];
        dhi[p] = hi[i];
        dy[p] = y[i];
    }
}

__global__ void k_wide_check_sorted(const uint64_t* __restrict__ lo, const uint64_t* __restrict__ hi, size_t n,
                                    uint32_t* flags) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x + 1; i < n; i += (size_t)gridDim.x * blockDim.x)
        if (!key_less(lo[i - 1], hi[i - 1], lo[i], hi[i])) atomicOr(flags, FLAG_BAD_INPUT);
}

// grouping hash of a key: its top `bits` bits (records are grouped by it before the
// exact merge; 48 by default)
__host__ __device__ __forceinline__ uint64_t group_hash(uint64_t lo, uint64_t hi, int bits) {
    return wide_hash(lo, hi) >> (64 - bits);
}

__global__ void k_wide_group_key(const uint64_t* __restrict__ lo, const uint64_t* __restrict__ hi, size_t n,
                                 int bits, uint64_t* __restrict__ key) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        key[i] = group_hash(lo[i], hi[i], bits);
}

__global__ void k_wide_run_heads(const uint64_t* __restrict__ h, size_t n, uint32_t* __restrict__ head) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        head[i] = (i == 0 || h[i] != h[i - 1]) ? 1u : 0u;
}

// Distinct keys of each run of equal grouping hash: 1 unless two keys share the hash.
__global__ void k_wide_run_count(const uint64_t* __restrict__ lo, const uint64_t* __restrict__ hi,
                                 const uint64_t* __restrict__ start, size_t n_run, uint32_t* __restrict__ cnt) {
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_run; r += (size_t)gridDim.x * blockDim.x) {
        const uint64_t b = start[r], e = start[r + 1];
        bool same = true;
        for (uint64_t i = b + 1; i < e && same; ++i) same = lo[i] == lo[b] && hi[i] == hi[b];
        uint32_t c = 1;
        if (!same) {
            c = 0;
            for (uint64_t i = b; i < e; ++i) {
                bool first = true;
                for (uint64_t q = b; q < i && first; ++q) first = !(lo[q] == lo[i] && hi[q] == hi[i]);
                c += first ? 1u : 0u;
            }
        }
        cnt[r] = c;
    }
}

// Exact sum of every distinct key of each run, written from opos[r] on (first occurrence
// order inside a run); the grouping hash goes along for the diag search.
__global__ void k_wide_run_sum(const uint64_t* __restrict__ lo, const uint64_t* __restrict__ hi,
                               const double* __restrict__ v, const uint64_t* __restrict__ h,
                               const uint64_t* __restrict__ start, const uint64_t* __restrict__ opos, size_t n_run,
                               uint64_t* __restrict__ olo, uint64_t* __restrict__ ohi, double* __restrict__ oy,
                               uint64_t* __restrict__ oh) {
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_run; r += (size_t)gridDim.x * blockDim.x) {
        const uint64_t b = start[r], e = start[r + 1];
        uint64_t o = opos[r];
        for (uint64_t i = b; i < e; ++i) {
            bool first = true;
            for (uint64_t q = b; q < i && first; ++q) first = !(lo[q] == lo[i] && hi[q] == hi[i]);
            if (!first) continue;
            fx128 acc = fx_zero();
            for (uint64_t q = i; q < e; ++q)
                if (lo[q] == lo[i] && hi[q] == hi[i]) acc = fx_add(acc, fx_from(v[q]));
            olo[o] = lo[i];
            ohi[o] = hi[i];
            oy[o] = fx_to(acc);
            oh[o] = h[b];
            ++o;
        }
    }
}

// diag += y_t c_t for the sources t of this pass; the merged list is ordered by grouping
// hash (every source emits its diagonal record, so each finds its own key)
__global__ __launch_bounds__(W_BLOCK) void k_wide_diag_h(const uint64_t* __restrict__ slo,
                                                         const uint64_t* __restrict__ shi,
                                                         const double* __restrict__ sv, size_t ns, uint32_t n_pass,
                                                         uint32_t pass, const uint64_t* __restrict__ ylo,
                                                         const uint64_t* __restrict__ yhi,
                                                         const double* __restrict__ y,
                                                         const uint64_t* __restrict__ yh, size_t n_y, int bits,
                                                         unsigned long long* acc, uint32_t* flags) {
    fx128 mine = fx_zero();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t lo = slo[i], hi = shi[i];
        if (wide_class(lo, hi, n_pass) != pass) continue;
        const uint64_t h = group_hash(lo, hi, bits);
        size_t a = 0, b = n_y;
        while (a < b) {
            const size_t m = (a + b) >> 1;
            if (yh[m] < h) a = m + 1;
            else b = m;
        }
        bool found = false;
        for (; a < n_y && yh[a] == h; ++a)
            if (ylo[a] == lo && yhi[a] == hi) {
                mine = fx_add(mine, fx_from(y[a] * sv[i]));
                found = true;
                break;
            }
        if (!found) atomicOr(flags, FLAG_BAD_INPUT);
    }
    fx_block_reduce_atomic<W_BLOCK>(mine, acc);
}

// ---- vector operations on two-word keys ----------------------------------------------
__global__ void k_wide_keyheads(const uint64_t* __restrict__ lo, const uint64_t* __restrict__ hi, size_t n,
                                uint32_t* __restrict__ head) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        head[i] = (i == 0 || lo[i] != lo[i - 1] || hi[i] != hi[i - 1]) ? 1u : 0u;
}

// One value per distinct key of a key-sorted list: the exact fixed-point sum of its entries
// (EXACT), or the plain double sum of its (at most two) entries.
template <bool EXACT>
__global__ void k_wide_keymerge(const uint64_t* __restrict__ lo, const uint64_t* __restrict__ hi,
                                const double* __restrict__ v, const uint64_t* __restrict__ start, size_t n_seg,
                                uint64_t* __restrict__ olo, uint64_t* __restrict__ ohi, double* __restrict__ ov,
                                uint32_t* flags) {
    bool bad = false;
    for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < n_seg; s += (size_t)gridDim.x * blockDim.x) {
        const uint64_t b = start[s], e = start[s + 1];
        double y;
        if (EXACT) {
            fx128 acc = fx_zero();
            for (uint64_t j = b; j < e; ++j) {
                bad |= !fx_in_range(v[j]);
                acc = fx_add(acc, fx_from(v[j]));
            }
            y = fx_to(acc);
        } else {
            y = v[b];
            for (uint64_t j = b + 1; j < e; ++j) y = y + v[j];
        }
        olo[s] = lo[b];
        ohi[s] = hi[b];
        ov[s] = y;
    }
    if (bad) atomicOr(flags, FLAG_RANGE);
}

// dst[off + i] = (key i of a two-word-key vector, a * value i)
__global__ void k_wide_scaled_split(const uint64_t* __restrict__ k2, const double* __restrict__ v, size_t n, double a,
                                    uint64_t* __restrict__ lo, uint64_t* __restrict__ hi, double* __restrict__ vo) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        lo[i] = k2[2 * i];
        hi[i] = k2[2 * i + 1];
        vo[i] = a * v[i];
    }
}

// exact sum of a_i b_t over the keys the two (key-sorted) vectors share
__global__ __launch_bounds__(W_BLOCK) void k_wide_dot(const uint64_t* __restrict__ ak2, const double* __restrict__ av,
                                                      size_t na, const uint64_t* __restrict__ bk2,
                                                      const double* __restrict__ bv, size_t nb,
                                                      unsigned long long* acc, uint32_t* flags) {
    fx128 mine = fx_zero();
    bool bad = false;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < na; i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t lo = ak2[2 * i], hi = ak2[2 * i + 1];
        size_t a = 0, b = nb;
        while (a < b) {
            const size_t m = (a + b) >> 1;
            if (key_less(bk2[2 * m], bk2[2 * m + 1], lo, hi)) a = m + 1;
            else b = m;
        }
        if (a < nb && bk2[2 * a] == lo && bk2[2 * a + 1] == hi) {
            const double p = av[i] * bv[a];
            bad |= !fx_in_range(p);
            mine = fx_add(mine, fx_from(p));
        }
    }
    if (bad) atomicOr(flags, FLAG_RANGE);
    fx_block_reduce_atomic<W_BLOCK>(mine, acc);
}

int grid_of_w(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 256 * 16)); }


uint64_t read_u64(Ctx* c, const uint64_t* d) {
    uint64_t h = 0;
    HIPCHK(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return h;
}

// Scratch of the wide path: thirteen 8-byte arrays of n + 1 entries.
struct WBufs {
    uint64_t* a[11];
    uint64_t* pos;
    uint64_t* start;
};

WBufs wide_bufs(Ctx* c, size_t n) {
    WBufs w;
    DevBuf* bufs[11] = {&c->cand_k[0], &c->cand_k[1], &c->cand_v[0], &c->cand_v[1], &c->tmp_k[0], &c->tmp_k[1],
                        &c->tmp_v[0],  &c->tmp_v[1],  &c->plan_first, &c->plan_str, &c->plan_cnt};
    for (int i = 0; i < 11; ++i) w.a[i] = bufs[i]->get<uint64_t>(n + 1);
    w.pos = c->src_off.get<uint64_t>(n + 1);
    w.start = c->part_base.get<uint64_t>(n + 1);
    return w;
}

// Sorts (lo, hi, v) by the 128-bit key hi:lo into (olo, ohi, ov); stable.  The four
// scratch arrays t0..t3 hold n entries each; none of them aliases an input or output.
void sort128(Ctx* c, const uint64_t* lo, const uint64_t* hi, const double* v, size_t n, int n_so, uint64_t* t0,
             uint64_t* t1, uint64_t* t2, uint64_t* t3, uint64_t* olo, uint64_t* ohi, double* ov) {
    if (n == 0) return;
    const int g = grid_of_w(n);
    HIPCHK(hipMemcpyAsync(t0, lo, n * 8, hipMemcpyDeviceToDevice, c->stream));
    k_wide_iota<<<g, 256, 0, c->stream>>>(t1, n);
    uint64_t* K = t0;
    uint64_t* I = t1;
    radix_sort_kv<uint64_t>(c, &K, &I, t2, t3, n, std::min(64, n_so));
    if (n_so > 64) {
        uint64_t* Kf = (K == t0) ? t2 : t0;
        uint64_t* If = (I == t1) ? t3 : t1;
        k_wide_gather<uint64_t><<<g, 256, 0, c->stream>>>(hi, I, n, Kf);
        uint64_t* K2 = Kf;
        uint64_t* I2 = I;
        radix_sort_kv<uint64_t>(c, &K2, &I2, K, If, n, n_so - 64);
        I = I2;
    }
    k_wide_gather<uint64_t><<<g, 256, 0, c->stream>>>(lo, I, n, olo);
    k_wide_gather<uint64_t><<<g, 256, 0, c->stream>>>(hi, I, n, ohi);
    k_wide_gather<double><<<g, 256, 0, c->stream>>>(v, I, n, ov);
    HIPCHK(hipGetLastError());
}

// Keeps the cap entries ranked highest by (|v| desc, key asc) of the key-sorted (lo, hi,
// v), in key order, in (olo, ohi, ov).  Returns the kept count.
size_t select128(Ctx* c, const uint64_t* lo, const uint64_t* hi, const double* v, size_t n, size_t cap, uint64_t* t0,
                 uint64_t* t1, uint64_t* t2, uint64_t* t3, uint64_t* pos, uint64_t* olo, uint64_t* ohi,
                 double* ov) {
    const int g = grid_of_w(n);
    const size_t keep = std::min(n, cap);
    uint32_t* fl = nullptr;
    if (n > cap) {
        k_wide_rank_key<<<g, 256, 0, c->stream>>>(v, n, t0);
        k_wide_iota<<<g, 256, 0, c->stream>>>(t1, n);
        uint64_t* K = t0;
        uint64_t* I = t1;
        radix_sort_kv<uint64_t>(c, &K, &I, t2, t3, n, 63);
        // the sorted index array is I (t1 or t3); the flags take whichever of t0/t2 is free
        fl = (uint32_t*)((K == t0) ? t2 : t0);
        HIPCHK(hipMemsetAsync(fl, 0, n * 4, c->stream));
        k_wide_mark<<<grid_of_w(cap), 256, 0, c->stream>>>(I, cap, fl);
    } else {
        HIPCHK(hipMemcpyAsync(olo, lo, n * 8, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(ohi, hi, n * 8, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(ov, v, n * 8, hipMemcpyDeviceToDevice, c->stream));
        return n;
    }
    exclusive_scan_u32(c, fl, n, pos, pos + n);
    k_wide_compact<<<g, 256, 0, c->stream>>>(fl, pos, n, lo, hi, v, olo, ohi, ov);
    HIPCHK(hipGetLastError());
    return keep;
}

// Growable staging list (lo, hi, y) that keeps its contents when it grows.
struct WStage {
    DevBuf *blo, *bhi, *by;
    size_t n = 0, cap = 0;
    uint64_t *lo = nullptr, *hi = nullptr;
    double* y = nullptr;
};

void stage_reserve(Ctx* c, WStage& s, size_t need, DevBuf& tmp) {
    if (need <= s.cap) return;
    const size_t ncap = std::max(need, s.cap * 2);
    DevBuf* bs[3] = {s.blo, s.bhi, s.by};
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < 3; ++i) {
        if (s.n) {
            uint64_t* t = tmp.get<uint64_t>(s.n);
            HIPCHK(hipMemcpyAsync(t, bs[i]->get<uint64_t>(s.n), s.n * 8, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            uint64_t* d = bs[i]->get<uint64_t>(ncap);
            HIPCHK(hipMemcpyAsync(d, t, s.n * 8, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        } else {
            bs[i]->get<uint64_t>(ncap);
        }
    }
    s.lo = s.blo->get<uint64_t>(ncap);
    s.hi = s.bhi->get<uint64_t>(ncap);
    s.y = s.by->get<double>(ncap);
    s.cap = ncap;
}

template <bool MIXED>
size_t wide_hv_t(Ctx* c, int n_so, const uint64_t* ik2, const double* iv, size_t n_in, size_t nd, double lambda,
                 double coeff, size_t out_cap, uint64_t* ok2, double* ov, double* diag) {
    const HamDev& H = c->ham;
    auto& st = c->st;
    kt_resolve(c);   // laps left pending by an earlier call that threw
    st = {};
    *diag = 0.0;
    PhaseTimer total(c), tm(c);
    if (n_in == 0) return 0;
    (void)flags_read_clear(c);
    // 1. input (key order checked), top-nd
    WBufs w = wide_bufs(c, n_in);
    uint64_t* in_lo = c->src_k.get<uint64_t>(n_in);
    uint64_t* in_hi = c->plan_k[0].get<uint64_t>(n_in);
    double* in_v = c->src_v.get<double>(n_in);
    k_wide_split<<<grid_of_w(n_in), 256, 0, c->stream>>>(ik2, iv, n_in, in_lo, in_hi, in_v);
    k_wide_check_sorted<<<grid_of_w(n_in), 256, 0, c->stream>>>(in_lo, in_hi, n_in, c->d_flags);
    if (flags_read_clear(c) & FLAG_BAD_INPUT) throw std::invalid_argument("input keys are not strictly ascending");
    const uint64_t* slo = in_lo;
    const uint64_t* shi = in_hi;
    const double* sv = in_v;
    size_t ns = n_in;
    if (n_in > nd) {
        uint64_t* klo = c->plan_k[1].get<uint64_t>(nd);
        uint64_t* khi = c->plan_v[0].get<uint64_t>(nd);
        double* kv = c->plan_v[1].get<double>(nd);
        ns = select128(c, in_lo, in_hi, in_v, n_in, nd, w.a[0], w.a[1], w.a[2], w.a[3], w.pos, klo, khi, kv);
        slo = klo;
        shi = khi;
        sv = kv;
    }
    st.n_src = ns;
    tm.lap(&st.ms_select);
    if (ns == 0) {
        kt_resolve(c);
        return 0;
    }
    // 2. passes: thirteen 8-byte scratch arrays per candidate record
    const uint64_t n_conn = H.n_conn;
    const uint64_t total_cand = (uint64_t)ns * n_conn;
    const size_t per_rec = 13 * 8;
    const size_t rec_cap = std::max<size_t>(c->ws_limit / per_rec, (size_t)1 << 16);
    uint32_t n_pass = (uint32_t)std::max<uint64_t>(1, (uint64_t)((double)total_cand * 1.05 / (double)rec_cap) + 1);
    if (total_cand <= rec_cap) n_pass = 1;
    WStage stg;
    stg.blo = &c->stage_k;
    stg.bhi = &c->lin_k[0];
    stg.by = &c->stage_v;
    fx128 dsum = fx_zero();
    for (int attempt = 0;; ++attempt) {
        stg.n = 0;
        dsum = fx_zero();
        st.n_candidates = 0;
        st.n_unique = 0;
        bool overflow = false;
        const size_t cap = std::min<size_t>(rec_cap, (size_t)(total_cand / n_pass + total_cand / (8 * n_pass) + 65536));
        w = wide_bufs(c, cap);
        for (uint32_t p = 0; p < n_pass && !overflow; ++p) {
            // generate
            HIPCHK(hipMemsetAsync(c->d_counter, 0, 8, c->stream));
            uint64_t *rlo = w.a[0], *rhi = w.a[1];
            double* rv = (double*)w.a[2];
            const int gg = (int)std::min<size_t>(ns, 256 * 8);
            k_wide_generate<MIXED><<<gg, W_BLOCK, 0, c->stream>>>(H, n_so, slo, shi, sv, ns, lambda, coeff, n_pass, p,
                                                                 rlo, rhi, rv, cap,
                                                                 (unsigned long long*)c->d_counter, c->d_flags);
            HIPCHK(hipGetLastError());
            const uint64_t n = n_pass == 1 ? total_cand : read_u64(c, c->d_counter);
            HIPCHK(hipStreamSynchronize(c->stream));
            tm.lap(&st.ms_generate);
            const uint32_t f = flags_read_clear(c);
            if (f & FLAG_RANGE) throw HipError("a term left the fixed-point range (|x| >= 2^42)", SFCI_E_RANGE);
            if (f & FLAG_BAD_INPUT)
                throw std::invalid_argument("an input determinant does not have n_elec/2 electrons of each spin "
                                            "inside n_so");
            if ((f & FLAG_GEN_OVERFLOW) || n > cap) {
                overflow = true;
                break;
            }
            st.n_candidates += n;
            if (n == 0) continue;
            // group by the key hash: stable LSD sort of (hash, index), then gathers
            const int g = grid_of_w(n);
            const int gbits = c->wide_group_bits;
            k_wide_group_key<<<g, 256, 0, c->stream>>>(rlo, rhi, n, gbits, w.a[3]);
            k_wide_iota<<<g, 256, 0, c->stream>>>(w.a[4], n);
            uint64_t* K = w.a[3];
            uint64_t* I = w.a[4];
            radix_sort_kv<uint64_t>(c, &K, &I, w.a[5], w.a[6], n, gbits);
            uint64_t *olo = w.a[7], *ohi = w.a[8];
            double* ovv = (double*)w.a[9];
            k_wide_gather<uint64_t><<<g, 256, 0, c->stream>>>(rlo, I, n, olo);
            k_wide_gather<uint64_t><<<g, 256, 0, c->stream>>>(rhi, I, n, ohi);
            k_wide_gather<double><<<g, 256, 0, c->stream>>>(rv, I, n, ovv);
            HIPCHK(hipGetLastError());
            tm.lap(&st.ms_sort);
            // runs of equal hash, distinct keys per run, exact sums (into the record arrays)
            uint64_t* Kfree = (K == w.a[3]) ? w.a[5] : w.a[3];
            uint64_t* Ifree = (I == w.a[4]) ? w.a[6] : w.a[4];
            uint32_t* head = (uint32_t*)Kfree;
            k_wide_run_heads<<<g, 256, 0, c->stream>>>(K, n, head);
            exclusive_scan_u32(c, head, n, w.pos, w.pos + n);
            const uint64_t n_run = read_u64(c, w.pos + n);
            k_wide_seg_start<<<g, 256, 0, c->stream>>>(head, w.pos, n, w.start);
            uint32_t* cnt = (uint32_t*)Ifree;
            k_wide_run_count<<<grid_of_w(n_run), 256, 0, c->stream>>>(olo, ohi, w.start, n_run, cnt);
            exclusive_scan_u32(c, cnt, n_run, w.a[10], w.a[10] + n_run);
            const uint64_t n_seg = read_u64(c, w.a[10] + n_run);
            uint64_t *ylo = rlo, *yhi = rhi;
            double* y = rv;
            uint64_t* yh = I;   // the sort's index array is free after the gathers
            k_wide_run_sum<<<grid_of_w(n_run), 256, 0, c->stream>>>(olo, ohi, ovv, K, w.start, w.a[10], n_run, ylo,
                                                                    yhi, y, yh);
            HIPCHK(hipGetLastError());
            st.n_unique += n_seg;
            // diag over the sources of this pass
            acc_clear(c, 0);
            k_wide_diag_h<<<grid_of_w(ns), W_BLOCK, 0, c->stream>>>(slo, shi, sv, ns, n_pass, p, ylo, yhi, y, yh,
                                                                   n_seg, gbits, c->d_acc, c->d_flags);
            HIPCHK(hipGetLastError());
            dsum = fx_add(dsum, acc_read_raw(c, 0));
            if (flags_read_clear(c) & FLAG_BAD_INPUT)
                throw HipError("a source's diagonal target was not found after the merge", SFCI_E_OVERFLOW);
            // cut, append to staging
            uint32_t* fl = (uint32_t*)Kfree;
            k_wide_flag_above<<<grid_of_w(n_seg), 256, 0, c->stream>>>(y, n_seg, c->cut, fl);
            exclusive_scan_u32(c, fl, n_seg, w.pos, w.pos + n_seg);
            const uint64_t n_keep = read_u64(c, w.pos + n_seg);
            stage_reserve(c, stg, stg.n + n_keep, c->lin_v[0]);
            k_wide_compact<<<grid_of_w(n_seg), 256, 0, c->stream>>>(fl, w.pos, n_seg, ylo, yhi, y, stg.lo + stg.n,
                                                                  stg.hi + stg.n, stg.y + stg.n);
            HIPCHK(hipGetLastError());
            stg.n += n_keep;
            tm.lap(&st.ms_reduce);
        }
        if (!overflow) break;
        if (attempt >= 6) throw HipError("candidate passes kept overflowing the workspace", SFCI_E_OVERFLOW);
        n_pass *= 2;
    }
    st.n_passes = n_pass;
    *diag = fx_to(dsum);
    // 4. output: key order, top out_cap
    const size_t m = stg.n;
    size_t n_out = 0;
    if (m) {
        WBufs w2 = wide_bufs(c, m);
        const uint64_t* flo = stg.lo;
        const uint64_t* fhi = stg.hi;
        const double* fy = stg.y;
        sort128(c, stg.lo, stg.hi, stg.y, m, n_so, w2.a[0], w2.a[1], w2.a[2], w2.a[3], w2.a[4], w2.a[5],
                (double*)w2.a[6]);
        flo = w2.a[4];
        fhi = w2.a[5];
        fy = (double*)w2.a[6];
        uint64_t *glo = w2.a[7], *ghi = w2.a[8];
        double* gy = (double*)w2.a[9];
        n_out = select128(c, flo, fhi, fy, m, out_cap, w2.a[0], w2.a[1], w2.a[2], w2.a[3], w2.pos, glo, ghi, gy);
        k_wide_join<<<grid_of_w(n_out), 256, 0, c->stream>>>(glo, ghi, gy, n_out, ok2, ov);
        HIPCHK(hipGetLastError());
    }
    tm.lap(&st.ms_select);
    total.lap(&st.ms_total);
    kt_resolve(c);   // synchronizes; folds the kernel and phase timings
    st.n_out = n_out;
    return n_out;
}

}  // namespace

size_t dev_hv_wide(Ctx* c, int n_so, const uint64_t* ik2, const double* iv, size_t n_in, size_t nd, double lambda,
                   double coeff, size_t out_cap, uint64_t* ok2, double* ov, double* diag) {
    return c->precision ? wide_hv_t<true>(c, n_so, ik2, iv, n_in, nd, lambda, coeff, out_cap, ok2, ov, diag)
                        : wide_hv_t<false>(c, n_so, ik2, iv, n_in, nd, lambda, coeff, out_cap, ok2, ov, diag);
}


// Key-sorted distinct keys of the concatenated list (lo, hi, v)[n] (n <= the WBufs size):
// sort, run heads, merged values into (olo, ohi, ov).  Uses a[0..7], pos and start.
static size_t wide_merge(Ctx* c, const WBufs& w, const uint64_t* lo, const uint64_t* hi, const double* v, size_t n,
                         bool exact, uint64_t* olo, uint64_t* ohi, double* ov) {
    if (n == 0) return 0;
    sort128(c, lo, hi, v, n, c->n_so, w.a[0], w.a[1], w.a[2], w.a[3], w.a[4], w.a[5], (double*)w.a[6]);
    uint32_t* head = (uint32_t*)w.a[7];
    const int g = grid_of_w(n);
    k_wide_keyheads<<<g, 256, 0, c->stream>>>(w.a[4], w.a[5], n, head);
    exclusive_scan_u32(c, head, n, w.pos, w.pos + n);
    const uint64_t n_seg = read_u64(c, w.pos + n);
    k_wide_seg_start<<<g, 256, 0, c->stream>>>(head, w.pos, n, w.start);
    if (exact)
        k_wide_keymerge<true><<<grid_of_w(n_seg), 256, 0, c->stream>>>(w.a[4], w.a[5], (double*)w.a[6], w.start,
                                                                       n_seg, olo, ohi, ov, c->d_flags);
    else
        k_wide_keymerge<false><<<grid_of_w(n_seg), 256, 0, c->stream>>>(w.a[4], w.a[5], (double*)w.a[6], w.start,
                                                                        n_seg, olo, ohi, ov, c->d_flags);
    HIPCHK(hipGetLastError());
    if (flags_read_clear(c) & FLAG_RANGE) throw HipError("a term left the fixed-point range (|x| >= 2^42)", SFCI_E_RANGE);
    return n_seg;
}

double dev_wide_dot(Ctx* c, const uint64_t* ak2, const double* av, size_t na, const uint64_t* bk2, const double* bv,
                    size_t nb) {
    acc_clear(c, 1);
    if (na && nb)
        k_wide_dot<<<grid_of_w(na), W_BLOCK, 0, c->stream>>>(ak2, av, na, bk2, bv, nb, c->d_acc + 2, c->d_flags);
    HIPCHK(hipGetLastError());
    const double d = acc_read(c, 1);
    if (flags_read_clear(c) & FLAG_RANGE) throw HipError("dot: term outside the fixed-point range", SFCI_E_RANGE);
    return d;
}

size_t dev_wide_add_two(Ctx* c, const uint64_t* vk2, const double* vv, size_t nv, const uint64_t* wk2,
                        const double* wv, size_t nw, double a, size_t cap, uint64_t* ok2, double* ov) {
    const size_t n = nv + nw;
    if (n == 0) return 0;
    WBufs w = wide_bufs(c, n);
    uint64_t *clo = w.a[8], *chi = w.a[9];
    double* cv = (double*)w.a[10];
    if (nv) k_wide_scaled_split<<<grid_of_w(nv), 256, 0, c->stream>>>(vk2, vv, nv, 1.0, clo, chi, cv);
    if (nw) k_wide_scaled_split<<<grid_of_w(nw), 256, 0, c->stream>>>(wk2, wv, nw, a, clo + nv, chi + nv, cv + nv);
    // the scan helpers need a[7]/pos/start only after the sort has consumed the list
    uint64_t *mlo = w.a[8], *mhi = w.a[9];
    double* mv = (double*)w.a[10];
    const size_t nm = wide_merge(c, w, clo, chi, cv, n, false, mlo, mhi, mv);
    const size_t n_out = select128(c, mlo, mhi, mv, nm, cap, w.a[0], w.a[1], w.a[2], w.a[3], w.pos, w.a[4], w.a[5],
                                   (double*)w.a[6]);
    k_wide_join<<<grid_of_w(n_out), 256, 0, c->stream>>>(w.a[4], w.a[5], (double*)w.a[6], n_out, ok2, ov);
    HIPCHK(hipGetLastError());
    return n_out;
}

size_t dev_wide_shrink(Ctx* c, const uint64_t* sk2, const double* sv, size_t n, size_t cap, uint64_t* ok2,
                       double* ov) {
    if (n == 0) return 0;
    WBufs w = wide_bufs(c, n);
    k_wide_scaled_split<<<grid_of_w(n), 256, 0, c->stream>>>(sk2, sv, n, 1.0, w.a[8], w.a[9], (double*)w.a[10]);
    const size_t n_out = select128(c, w.a[8], w.a[9], (double*)w.a[10], n, cap, w.a[0], w.a[1], w.a[2], w.a[3],
                                   w.pos, w.a[4], w.a[5], (double*)w.a[6]);
    k_wide_join<<<grid_of_w(n_out), 256, 0, c->stream>>>(w.a[4], w.a[5], (double*)w.a[6], n_out, ok2, ov);
    HIPCHK(hipGetLastError());
    return n_out;
}

size_t dev_wide_lincomb(Ctx* c, const sfci_wvec* const* vecs, const double* coef, int nvec, size_t cap, uint64_t* ok2,
                        double* ov) {
    size_t n = 0;
    for (int i = 0; i < nvec; ++i) n += (size_t)vecs[i]->len;
    if (n == 0) return 0;
    WBufs w = wide_bufs(c, n);
    size_t o = 0;
    for (int i = 0; i < nvec; ++i) {
        const size_t l = (size_t)vecs[i]->len;
        if (l)
            k_wide_scaled_split<<<grid_of_w(l), 256, 0, c->stream>>>(vecs[i]->k, vecs[i]->v, l, coef[i], w.a[8] + o,
                                                                     w.a[9] + o, (double*)w.a[10] + o);
        o += l;
    }
    const size_t nm = wide_merge(c, w, w.a[8], w.a[9], (double*)w.a[10], n, true, w.a[8], w.a[9], (double*)w.a[10]);
    const size_t n_out = select128(c, w.a[8], w.a[9], (double*)w.a[10], nm, cap, w.a[0], w.a[1], w.a[2], w.a[3],
                                   w.pos, w.a[4], w.a[5], (double*)w.a[6]);
    k_wide_join<<<grid_of_w(n_out), 256, 0, c->stream>>>(w.a[4], w.a[5], (double*)w.a[6], n_out, ok2, ov);
    HIPCHK(hipGetLastError());
    return n_out;
}


// <v|H|v>: the exact diag of H.v over all of v (every sum exact, one rounding per target
// value and one for the total; the 64-bit calc_energy pull rounds per source instead).
double dev_wide_energy(Ctx* c, const uint64_t* k2, const double* v, size_t n) {
    if (n == 0) return 0.0;
    uint64_t* ok2 = c->lin_v[1].get<uint64_t>(4);
    double* ov = (double*)(ok2 + 2);
    double d = 0.0;
    dev_hv_wide(c, c->n_so, k2, v, n, n, 0.0, 1.0, 1, ok2, ov, &d);
    return d;
}

}  // namespace sfci
