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
//   2. passes : targets are split into n_pass classes by the top word of a 64-bit hash of
//               the key, so that one class's candidates fit the workspace.
//   3. per pass
//      generate one workgroup per source: the diagonal, singles, same-spin and
//               opposite-spin doubles, each connection evaluated once with its full
//               Slater-Condon element (exact limb sums); a connection is kept when its
//               hash class is the pass, written as a 24-byte record (lo, hi, value) with
//               its 2-byte hash bucket (hb bits of the hash's low word); the diagonal
//               record carries the source index, its term is kept per source
//      partition two-level hash partition of the records by bucket (the 64-bit path's
//               compact-record partition: fixed-capacity regions by the top 8 bucket bits,
//               then each region by the rest)
//      reduce   one workgroup per bucket merges its records in an LDS open-addressing
//               table keyed by both words, with exact three-limb sums; a bucket with more
//               targets than the table holds is merged in hash classes, one class per
//               table fill; finished targets give diag (sources of this pass) and the
//               survivors |y| > cut, appended to the staging list
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

// Placement hash of a key: the XOR of a fixed pseudo-random 64-bit word per occupied spin
// orbital (a random linear map of the occupation bits), so that a connection's hash is its
// source's hash XOR the words of the 2 or 4 spin orbitals it changes.  The pass class is a
// range of its top word; the bucket is the top hb bits of its low word.
__host__ __device__ __forceinline__ uint64_t orb_word(int p) {
    uint64_t z = 0x9E3779B97F4A7C15ull * (uint64_t)(p + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
constexpr int W_ORBS = 128;
// placement hash of (lo, hi) from the word table t (used where no source hash is known)
__device__ __forceinline__ uint64_t place_hash(const uint64_t* t, uint64_t lo, uint64_t hi) {
    uint64_t h = 0;
    for (uint64_t b = lo; b; b &= b - 1) h ^= t[__ffsll((long long)b) - 1];
    for (uint64_t b = hi; b; b &= b - 1) h ^= t[64 + __ffsll((long long)b) - 1];
    return h;
}
__device__ __forceinline__ void load_orb_words(uint64_t* t) {
    for (int i = threadIdx.x; i < W_ORBS; i += blockDim.x) t[i] = orb_word(i);
}
__host__ __device__ __forceinline__ uint32_t pass_of(uint64_t h, uint32_t n_pass) {
    return (uint32_t)(((h >> 32) * (uint64_t)n_pass) >> 32);
}
__host__ __device__ __forceinline__ uint32_t bucket_of(uint64_t h, int hb) { return (uint32_t)h >> (32 - hb); }
// Inside a bucket the LDS-table slot is the top 16 bits of wide_hash's low word and the
// merge class of an overflowing bucket the low 16 bits of its top word (a hash independent
// of the placement hash).
__host__ __device__ __forceinline__ uint32_t slot16_of(uint64_t g) { return (uint32_t)g >> 16; }
__host__ __device__ __forceinline__ uint32_t class16_of(uint64_t g) { return (uint32_t)(g >> 32) & 0xffffu; }

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

// single excitations of one spin: k (n - k) <= 31 * 33 with fewer than 64 electrons
constexpr int W_SINGLES = 1024;

// Incremental quotient/remainder of i / d for i = start, start + step, ...
struct WDiv {
    uint32_t q, r, dq, dr, d;
    __device__ __forceinline__ WDiv(uint32_t start, uint32_t step, uint32_t div)
        : q(start / div), r(start % div), dq(step / div), dr(step % div), d(div) {}
    __device__ __forceinline__ void next() {
        q += dq;
        r += dr;
        if (r >= d) {
            r -= d;
            ++q;
        }
    }
};

struct WSrc {
    uint64_t lo, hi;
    uint64_t hx;        // placement hash
    double temp;        // coeff * c_j
    int n_occ_so;
    int ok;             // the determinant has k electrons of each spin inside n_so
    uint8_t occ_so[64]; // occupied spin orbitals, ascending
    uint8_t occ[2][32]; // occupied spatial orbitals of each spin, ascending
    uint8_t vir[2][64]; // empty spatial orbitals of each spin, ascending
    uint16_t sx[2][W_SINGLES];   // single x of each spin: occupied spin orbital | empty one << 8
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

// Passes generated together by one launch (each into its own region of the record buffer).
constexpr int W_BATCH = 32;   // target passes generated by one launch, at most

// Where a source's records go.  Single pass: record `idx` of source j at j * n_conn + idx
// (every connection is kept; no atomics).  Several passes: a launch generates the passes
// [p_lo, p_lo + nb), pass p_lo + r into region r (cap records at r * cap); each source's
// records of each region are counted first and reserved with one global atomic per region
// (rbase), and each record then takes its slot from the region's workgroup cursor in LDS
// with one LDS atomic (the order inside a region is free: every sum is exact).  Each record
// goes with its hash bucket; both are written with streaming stores (read back only by the
// partition).
struct WEmit {
    RecW* rec;
    uint16_t* bkt;
    int hb;
    uint64_t cap;                        // records per region
    uint64_t base;                       // single pass: j * n_conn
    const unsigned long long* rbase;     // LDS, per region
    unsigned* cursor;                    // LDS, per region; nullptr in single-pass mode
    uint32_t* flags;
    __device__ __forceinline__ void store(uint64_t at, uint64_t lo, uint64_t hi, double val, uint64_t h) const {
        __builtin_nontemporal_store(lo, &rec[at].lo);
        __builtin_nontemporal_store(hi, &rec[at].hi);
        __builtin_nontemporal_store(val, &rec[at].val);
        __builtin_nontemporal_store((uint16_t)bucket_of(h, hb), &bkt[at]);
    }
    // r is the record's region (-1: not in this launch's passes), h its placement hash
    __device__ __forceinline__ void put(int r, uint32_t idx, uint64_t lo, uint64_t hi, double val,
                                        uint64_t h) const {
        if (r < 0) return;
        if (!cursor) {
            const uint64_t i = base + idx;
            if (i < cap) store(i, lo, hi, val, h);
            else atomicOr(flags, FLAG_GEN_OVERFLOW);
            return;
        }
        const uint64_t i = rbase[r] + atomicAdd(&cursor[r], 1u);
        if (i < cap) store((uint64_t)r * cap + i, lo, hi, val, h);
        else atomicOr(flags, FLAG_GEN_OVERFLOW);
    }
};

// Stages determinant (lo, hi) in s: its occupied spin orbitals, and the occupied and empty
// spatial orbitals of each spin.  Run by one full wavefront; s.ok = 0 when the determinant
// does not have k electrons of each spin inside n_so (FLAG_BAD_INPUT is then raised).
__device__ __forceinline__ void stage_wsrc(const HamDev& H, int n_so, uint64_t lo, uint64_t hi, double temp,
                                           const uint64_t* t_orb, WSrc& s, uint32_t* flags) {
    const int lane = threadIdx.x & (SFCI_WAVE - 1);
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int N = H.n_orb, k = H.k;
    int n_occ = 0;
    bool ok = true;
    uint64_t hx = 0;
    for (int half = 0; half < 2; ++half) {
        const int p = half * 64 + lane;
        const bool oc = bit_of(lo, hi, p);
        if (oc) hx ^= t_orb[p];
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
    for (int off = SFCI_WAVE / 2; off > 0; off >>= 1) hx ^= __shfl_xor(hx, off, SFCI_WAVE);
    if (lane == 0) {
        s.lo = lo;
        s.hi = hi;
        s.hx = hx;
        s.temp = temp;
        s.n_occ_so = n_occ;
        s.ok = ok ? 1 : 0;
        if (!ok) atomicOr(flags, FLAG_BAD_INPUT);
    }
}

// E_core + sum_occ h_mm + sum_{m != p} 1/2 [(mm|pp) - (mp|pm)] of the staged source, summed
// exactly by one full wavefront; the rounded value is valid in lane 0, ok = every term in
// the fixed-point range (wave-uniform).
__device__ __forceinline__ double wide_diag_elem(const HamDev& H, const WSrc& s, bool& ok) {
    const int lane = threadIdx.x & (SFCI_WAVE - 1);
    const int n_occ = s.n_occ_so;
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
    ok = __all(dok);
    return fx_to(fx_add(fx_from(H.constant), f));
}

// <D'|H|D> for the single n -> m (spin orbitals) of the staged source: h_mn + sum over the
// other occupied p of (mn|pp) - (mp|pn), summed exactly; NaN when a term leaves the range.
__device__ __forceinline__ double single_elem_w(const HamDev& H, const WSrc& s, int n, int m) {
    FxLimbs acc = fxl_zero();
    bool ok = fxl_add(acc, h_w(H, m, n));
    for (int b = 0; b < s.n_occ_so; ++b) {
        const int p = s.occ_so[b];
        if (p == n) continue;
        ok &= fxl_add(acc, g_w(H, m, n, p, p));
        ok &= fxl_add(acc, -g_w(H, m, p, p, n));
    }
    return ok ? sign_between_w(s.lo, s.hi, m, n) * fx_to(fxl_join(acc)) : __builtin_nan("");
}

// The single-excitation table of the staged source (s.sx); run by the whole workgroup
// after stage_wsrc, followed by a barrier.
__device__ __forceinline__ void stage_singles(const HamDev& H, WSrc& s) {
    const uint32_t kv = (uint32_t)H.n_sa, v = (uint32_t)H.v;
    for (uint32_t e = threadIdx.x; e < 2 * kv; e += blockDim.x) {
        const uint32_t sp = e >= kv ? 1u : 0u, x = e - sp * kv;
        const uint32_t i = x / v, a = x - i * v;
        s.sx[sp][x] = (uint16_t)((2 * s.occ[sp][i] + sp) | ((2 * s.vir[sp][a] + sp) << 8));
    }
}

// Enumerates the off-diagonal connections of the staged source with this thread's share
// of the index range (stride W_BLOCK), calling f(e, tl, th, hx, n, q, x, y) with the
// target's key and placement hash and the spin orbitals of the excitation (singles: n -> x,
// q = y = -1; doubles: n < q -> x < y).  e is the connection's index (1-based; 0 is the
// diagonal): singles (spin 0, spin 1), same-spin doubles (spin 0, spin 1), opposite-spin
// doubles (spin-0 single * kv + spin-1 single).  t_orb: the placement words.
template <class F>
__device__ __forceinline__ void for_connections(const HamDev& H, const WSrc& s, const uint64_t* t_orb, F&& f) {
    const int tid = threadIdx.x;
    const uint32_t kv = (uint32_t)H.n_sa, dss = (uint32_t)H.n_daa;
    const uint64_t lo = s.lo, hi = s.hi, hs = s.hx;
    for (int sp = 0; sp < 2; ++sp) {
        for (uint32_t x = tid; x < kv; x += W_BLOCK) {
            const uint32_t sx = s.sx[sp][x];
            const int n = sx & 0xff, m = sx >> 8;
            uint64_t tl = lo, th = hi;
            flip(tl, th, n);
            flip(tl, th, m);
            f(1 + sp * kv + x, tl, th, hs ^ t_orb[n] ^ t_orb[m], n, -1, m, -1);
        }
    }
    for (int sp = 0; sp < 2 && dss; ++sp) {
        WDiv it(tid, W_BLOCK, H.n_vpair);
        for (uint32_t x = tid; x < dss; x += W_BLOCK, it.next()) {
            const uint16_t po = H.occ_pairs[it.q], pv = H.vir_pairs[it.r];
            const int n = 2 * s.occ[sp][po & 0xff] + sp, q = 2 * s.occ[sp][po >> 8] + sp;
            const int xx = 2 * s.vir[sp][pv & 0xff] + sp, y = 2 * s.vir[sp][pv >> 8] + sp;
            uint64_t tl = lo, th = hi;
            flip(tl, th, n);
            flip(tl, th, q);
            flip(tl, th, xx);
            flip(tl, th, y);
            f(1 + 2 * kv + sp * dss + x, tl, th, hs ^ t_orb[n] ^ t_orb[q] ^ t_orb[xx] ^ t_orb[y], n, q, xx, y);
        }
    }
    if (kv == 0) return;
    const uint32_t dos = kv * kv;
    WDiv it(tid, W_BLOCK, kv);
    for (uint32_t e = tid; e < dos; e += W_BLOCK, it.next()) {
        const uint32_t x0 = s.sx[0][it.q], x1 = s.sx[1][it.r];
        const int ni = x0 & 0xff, xa = x0 >> 8, nj = x1 & 0xff, yb = x1 >> 8;
        const int n = min(ni, nj), q = max(ni, nj), xx = min(xa, yb), y = max(xa, yb);
        uint64_t tl = lo, th = hi;
        flip(tl, th, n);
        flip(tl, th, q);
        flip(tl, th, xx);
        flip(tl, th, y);
        f(1 + 2 * kv + 2 * dss + e, tl, th, hs ^ t_orb[ni] ^ t_orb[xa] ^ t_orb[nj] ^ t_orb[yb], n, q, xx, y);
    }
}

// H_tj of a connection reported by for_connections
__device__ __forceinline__ double conn_elem(const HamDev& H, const WSrc& s, int n, int q, int x, int y) {
    return q < 0 ? single_elem_w(H, s, n, x) : double_elem_w(H, s.lo, s.hi, n, q, x, y);
}

// Connections of each source whose pass is in [p_lo, p_lo + nb) (n_pass == 1: all, at
// j * n_conn + e): (key, H_tj coeff c_j) for every off-diagonal connected t, and the
// diagonal record (the source's own key, tagged with its index j) with dterm[j] =
// coeff (H_jj - lambda) c_j.  count[r]: records of region r (several passes).
template <bool MIXED>
__global__ __launch_bounds__(W_BLOCK) void k_wide_generate(HamDev H, int n_so, const uint64_t* __restrict__ slo,
                                                           const uint64_t* __restrict__ shi,
                                                           const double* __restrict__ sv, size_t ns, double lambda,
                                                           double coeff, uint32_t n_pass, uint32_t p_lo, int nb,
                                                           RecW* __restrict__ rec, uint16_t* __restrict__ bkt, int hb,
                                                           uint64_t cap, unsigned long long* count,
                                                           double* __restrict__ dterm, uint32_t* flags,
                                                           unsigned long long* src_next) {
    __shared__ WSrc s;
    __shared__ unsigned s_cnt[W_BATCH], s_cur[W_BATCH];
    __shared__ unsigned long long s_base[W_BATCH];
    __shared__ uint64_t t_orb[W_ORBS];
    const int tid = threadIdx.x, lane = tid & (SFCI_WAVE - 1), wv = tid / SFCI_WAVE;
    load_orb_words(t_orb);
    // region of a placement hash in this launch, or -1
    auto region = [&](uint64_t h) -> int {
        const uint32_t p = pass_of(h, n_pass) - p_lo;   // wraps for p < p_lo
        return p < (uint32_t)nb ? (int)p : -1;
    };
    // sources from a global counter (a workgroup's first source is its index): the source after
    // next is claimed when one starts and published at its end
    __shared__ unsigned long long s_nj;
    unsigned long long nn = 0;
    if (tid == 0) s_nj = gridDim.x + atomicAdd(src_next, 1ull);
    __syncthreads();
    auto next_source = [&]() -> size_t {
        __syncthreads();
        const size_t jn = (size_t)s_nj;
        __syncthreads();   // every thread has read s_nj
        if (tid == 0) s_nj = nn;
        return jn;
    };
    for (size_t j = blockIdx.x; j < ns; j = next_source()) {
        if (tid == 0) nn = gridDim.x + atomicAdd(src_next, 1ull);
        __syncthreads();   // the previous source's lists are no longer read
        if (wv == 0) stage_wsrc(H, n_so, slo[j], shi[j], coeff * sv[j], t_orb, s, flags);
        __syncthreads();
        if (!s.ok) continue;   // block-uniform
        stage_singles(H, s);
        __syncthreads();
        const uint64_t lo = s.lo, hi = s.hi;
        const double temp = s.temp;
        const uint32_t n_conn = (uint32_t)H.n_conn;
        bool bad = false;
        WEmit em{rec, bkt, hb, cap, (uint64_t)j * n_conn, s_base, nullptr, flags};
        if (n_pass > 1) {
            // count this source's connections of each region (LDS atomics), reserve them with
            // one global atomic per region
            if (tid < W_BATCH) {
                s_cnt[tid] = 0;
                s_cur[tid] = 0;
            }
            __syncthreads();
            if (tid == 0) {
                const int r = region(s.hx);
                if (r >= 0) atomicAdd(&s_cnt[r], 1u);
            }
            for_connections(H, s, t_orb, [&](uint32_t, uint64_t, uint64_t, uint64_t hx, int, int, int, int) {
                const int r = region(hx);
                if (r >= 0) atomicAdd(&s_cnt[r], 1u);
            });
            __syncthreads();
            if (tid < nb) s_base[tid] = s_cnt[tid] ? atomicAdd(&count[tid], (unsigned long long)s_cnt[tid]) : 0ull;
            __syncthreads();
            em.cursor = s_cur;
        }
        // diagonal: its term per source, a tagged record for the merge
        if (wv == 0) {
            const uint64_t h = s.hx;
            const int r = region(h);
            if (r >= 0) {   // wave-uniform
                bool dok;
                const double e = wide_diag_elem(H, s, dok);
                if (lane == 0) {
                    const double dv = (e - lambda) * temp;
                    if (!dok || !fx_in_range(dv)) bad = true;
                    dterm[j] = dv;
                }
                em.put(lane == 0 ? r : -1, 0, lo, hi, wtag_val(j), h);
            }
        }
        for_connections(H, s, t_orb, [&](uint32_t e, uint64_t tl, uint64_t th, uint64_t h, int n, int q, int x, int y) {
            const int r = region(h);
            double val = 0.0;
            if (r >= 0) {
                val = wide_val<MIXED>(conn_elem(H, s, n, q, x, y), temp);
                bad |= !fx_in_range(val);
            }
            em.put(r, e, tl, th, val, h);
        });
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

__global__ void k_fill_u64(unsigned long long* __restrict__ x, size_t n, unsigned long long val) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] = val;
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

// ---- bucket merge ------------------------------------------------------------------
// One workgroup per bucket of partitioned records (WR_SLOTS-entry LDS table: both key
// words and three exact limbs per slot).  Same class-splitting scheme as the 64-bit
// merge: a class whose table passes fill_max occupied slots is split into subclasses
// (class cq of cm -> cq*f .. cq*f+f-1 of cm*f on the 16-bit class value) and each is merged
// by rereading the bucket; beyond WR_MAX_CLASSES classes FLAG_BUCKET_FULL is raised.
constexpr int WR_THREADS = 1024;
constexpr int WR_SLOTS = 3584;                 // 3584 x 40 B = 140 KB of LDS per workgroup
constexpr int WR_BATCH = 4;                    // records per thread loaded together
constexpr int WR_TAGS = 512;                   // tagged targets per class finished from a list
constexpr int WR_FILL = WR_SLOTS * 4 / 5;      // occupied slots at which a table counts as full
constexpr int WR_TARGET = WR_SLOTS * 3 / 5;    // targets per class aimed at when a bucket splits
constexpr int WR_STACK = 256;                  // pending classes of one bucket
constexpr uint32_t WR_MAX_CLASSES = 4096;
constexpr double WR_PASS_LOAD = 0.55;          // distinct targets of a pass per 2^16 tables, in slots
constexpr unsigned long long WR_EMPTY = ~0ull;   // no key word is all ones (n_elec < 64)
static_assert(WR_TAGS <= WR_THREADS, "the tag list is finished one entry per thread");

// index of key (lo, hi) in the key-ordered source list, or ns
__device__ __forceinline__ size_t find_source(const uint64_t* __restrict__ slo, const uint64_t* __restrict__ shi,
                                              size_t ns, uint64_t lo, uint64_t hi) {
    size_t a = 0, b = ns;
    while (a < b) {
        const size_t m = (a + b) >> 1;
        if (key_less(slo[m], shi[m], lo, hi)) a = m + 1;
        else b = m;
    }
    return (a < ns && slo[a] == lo && shi[a] == hi) ? a : ns;
}

// Exact merge of the records of each bucket [bstart[b], bend[b]) of a pass: y_t = the sum of
// the target's terms, plus dterm[j] for a source's own key (tagged record, index j), whose
// y_t c_j goes into diag_acc; |y_t| > cut is appended (lo, hi, y) from out_count on.
//
// A slot's key is claimed word by word with compare-and-swap on each word (no lock, no
// wait): a record takes the slot whose lo word is its lo (claiming an empty one), then the
// slot's hi word if it is empty; if the hi word holds another key's hi it probes on.  Both
// words never change once set, so every record of a key follows the same probe sequence
// to the same slot, and the claimer of a lo word always completes the slot's hi word (it
// claims it or finds it taken), so no slot is left half set.
__global__ __launch_bounds__(WR_THREADS) void k_wide_bucket_reduce(
    const RecW* __restrict__ rec, const uint64_t* __restrict__ bstart, const uint64_t* __restrict__ bend,
    uint64_t n_buckets, int hb, const uint64_t* __restrict__ slo, const uint64_t* __restrict__ shi,
    const double* __restrict__ sv, const double* __restrict__ dterm, size_t ns, double cut,
    uint64_t* __restrict__ olo, uint64_t* __restrict__ ohi, double* __restrict__ oy, unsigned long long* out_count,
    size_t out_cap, unsigned long long* unique_count, unsigned long long* diag_acc, uint32_t* flags,
    unsigned long long* split_count, uint32_t fill_max, unsigned long long* bucket_next) {
    __shared__ unsigned long long t_lo[WR_SLOTS], t_hi[WR_SLOTS];
    __shared__ unsigned long long t_l0[WR_SLOTS], t_l1[WR_SLOTS], t_l2[WR_SLOTS];
    __shared__ uint32_t t_tag[(WR_SLOTS + 31) / 32];   // tagged, finished by search
    __shared__ uint32_t t_tl[(WR_SLOTS + 31) / 32];    // tagged, finished from the tag list
    __shared__ uint32_t s_tslot[WR_TAGS], s_tj[WR_TAGS];
    __shared__ int s_full, s_ntag, s_sp, s_bad;
    __shared__ uint32_t s_fill, s_nread;
    __shared__ uint32_t s_cm[WR_STACK], s_cq[WR_STACK];   // pending classes: (count, index)
    __shared__ uint32_t s_wtot[WR_THREADS / SFCI_WAVE];
    __shared__ unsigned long long s_obase;
    const int tid = threadIdx.x, lane = tid & (SFCI_WAVE - 1);
    if (tid == 0) s_bad = 0;
    fx128 dacc = fx_zero();
    unsigned long long uniq = 0;
    bool bad_input = false, overflow = false, any_full = false;
    auto merge_bucket = [&](const uint64_t st, const uint64_t en) {
        if (tid == 0) {
            s_sp = 1;
            s_cm[0] = 1;
            s_cq[0] = 0;
        }
        for (;;) {
            __syncthreads();
            const int sp = s_sp;
            if (sp == 0) break;
            const uint32_t cm = s_cm[sp - 1], cq = s_cq[sp - 1];
            __syncthreads();
            if (tid == 0) s_sp = sp - 1;
            uint32_t n_ins = 0;
            constexpr uint64_t STEP = (uint64_t)WR_BATCH * WR_THREADS;
            auto load_full = [&](uint64_t i0, uint64_t* l_, uint64_t* h_, double* v_) {
#pragma unroll
                for (int r = 0; r < WR_BATCH; ++r) {
                    const RecW x = rec[i0 + (uint64_t)r * WR_THREADS + tid];
                    l_[r] = x.lo;
                    h_[r] = x.hi;
                    v_[r] = x.val;
                }
            };
            auto load_round = [&](uint64_t i0, uint64_t* l_, uint64_t* h_, double* v_) {
#pragma unroll
                for (int r = 0; r < WR_BATCH; ++r) {
                    const uint64_t i = i0 + (uint64_t)r * WR_THREADS + tid;
                    l_[r] = WR_EMPTY;
                    h_[r] = 0;
                    v_[r] = 0.0;
                    if (i < en) {
                        const RecW x = rec[i];
                        l_[r] = x.lo;
                        h_[r] = x.hi;
                        v_[r] = x.val;
                    }
                }
            };
            auto insert_round = [&](const uint64_t* l_, const uint64_t* h_, const double* v_) {
#pragma unroll
                for (int r = 0; r < WR_BATCH; ++r) {
                    const uint64_t lo = l_[r], hi = h_[r];
                    if (lo == WR_EMPTY) continue;
                    // a full table's class is merged again as subclasses: stop inserting
                    if (__hip_atomic_load(&s_full, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
                    const uint64_t g = wide_hash(lo, hi);
                    if ((uint32_t)((class16_of(g) * cm) >> 16) != cq) continue;
                    ++n_ins;
                    uint32_t slot = (slot16_of(g) * (uint32_t)WR_SLOTS) >> 16;
                    int probe = 0;
                    for (; probe < WR_SLOTS; ++probe) {
                        if ((probe & 63) == 63 &&
                            __hip_atomic_load(&s_full, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                            probe = WR_SLOTS;   // the class is being abandoned: give up the search
                            break;
                        }
                        unsigned long long cl =
                            __hip_atomic_load(&t_lo[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (cl == WR_EMPTY) {
                            const unsigned long long prev = atomicCAS(&t_lo[slot], WR_EMPTY, lo);
                            cl = prev == WR_EMPTY ? lo : prev;
                        }
                        if (cl == lo) {
                            unsigned long long ch =
                                __hip_atomic_load(&t_hi[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            if (ch == WR_EMPTY) {
                                const unsigned long long prev = atomicCAS(&t_hi[slot], WR_EMPTY, hi);
                                if (prev == WR_EMPTY) {
                                    if (atomicAdd(&s_fill, 1u) + 1u >= fill_max) s_full = 1;
                                    break;
                                }
                                ch = prev;
                            }
                            if (ch == hi) break;
                        }
                        slot = (slot + 1 == (uint32_t)WR_SLOTS) ? 0u : slot + 1;
                    }
                    if (probe == WR_SLOTS) {
                        s_full = 1;
                        continue;
                    }
                    const uint64_t j = wtag_index(v_[r]);
                    if (j != ~0ull) {
                        // the target's own source: (slot, index) into the tag list, or, when
                        // that is full, a tag bit for a search at the finish
                        const int t = atomicAdd(&s_ntag, 1);
                        if (t < WR_TAGS) {
                            s_tslot[t] = slot;
                            s_tj[t] = (uint32_t)j;
                            atomicOr(&t_tl[slot >> 5], 1u << (slot & 31));
                        } else {
                            atomicOr(&t_tag[slot >> 5], 1u << (slot & 31));
                        }
                        continue;
                    }
                    unsigned long long l0, l1, l2;
                    limb_split(v_[r], l0, l1, l2);
                    if (l0) atomicAdd(&t_l0[slot], l0);
                    if (l1) atomicAdd(&t_l1[slot], l1);
                    if (l2) atomicAdd(&t_l2[slot], l2);
                }
            };
            // the first round's loads are issued before the table is cleared
            uint64_t la[WR_BATCH], ha[WR_BATCH];
            double va[WR_BATCH];
            if (st + STEP <= en) load_full(st, la, ha, va);
            else load_round(st, la, ha, va);
            for (int i = tid; i < WR_SLOTS; i += WR_THREADS) {
                t_lo[i] = WR_EMPTY;
                t_hi[i] = WR_EMPTY;
                t_l0[i] = 0;
                t_l1[i] = 0;
                t_l2[i] = 0;
            }
            for (int i = tid; i < (WR_SLOTS + 31) / 32; i += WR_THREADS) {
                t_tag[i] = 0;
                t_tl[i] = 0;
            }
            if (tid == 0) {
                s_full = 0;
                s_fill = 0;
                s_ntag = 0;
                s_nread = 0;
            }
            __syncthreads();
            insert_round(la, ha, va);
            uint64_t i0 = st + STEP;
            // a full table ends the class early: each thread stops at its next round
            for (; i0 + STEP <= en && !__hip_atomic_load(&s_full, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                 i0 += STEP) {
                load_full(i0, la, ha, va);
                insert_round(la, ha, va);
            }
            if (i0 < en && !__hip_atomic_load(&s_full, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                load_round(i0, la, ha, va);
                insert_round(la, ha, va);
            }
            __syncthreads();
            if (s_full != 0) {
                // split this class into f subclasses, f from the table's fill after n_read of
                // its records
                uint32_t r = n_ins;
                for (int o = SFCI_WAVE / 2; o > 0; o >>= 1) r += __shfl_down(r, o, SFCI_WAVE);
                if (lane == 0 && r) atomicAdd(&s_nread, r);
                __syncthreads();
                if (tid == 0) {
                    uint32_t f = split_count_for(s_fill, s_nread, WR_TARGET, WR_STACK);
                    if (f < 2) f = 2;
                    if (s_sp + (int)f > WR_STACK) f = (uint32_t)(WR_STACK - s_sp);
                    if (f < 2 || cm * f > WR_MAX_CLASSES) {
                        s_bad = 1;
                    } else {
                        for (uint32_t i = 0; i < f; ++i) {
                            s_cm[s_sp + i] = cm * f;
                            s_cq[s_sp + i] = cq * f + i;
                        }
                        s_sp += (int)f;
                        atomicAdd(split_count, 1ull);
                    }
                }
                __syncthreads();
                if (s_bad) {
                    any_full = true;
                    if (tid == 0) s_sp = 0;
                }
                continue;
            }
            // finish every slot's target and the tag list's; the kept ones are written after
            // one block-wide scan of the per-thread counts, one global reservation per class
            constexpr int FR = (WR_SLOTS + WR_THREADS - 1) / WR_THREADS;
            bool em[FR + 1];
            double yv[FR + 1];
            uint64_t kl[FR + 1], kh[FR + 1];
            uint32_t ne = 0;
#pragma unroll
            for (int r = 0; r < FR; ++r) {
                const int slot = r * WR_THREADS + tid;
                em[r] = false;
                yv[r] = 0.0;
                kl[r] = 0;
                kh[r] = 0;
                const unsigned long long hw = slot < WR_SLOTS ? t_hi[slot] : WR_EMPTY;
                if (hw != WR_EMPTY) {
                    ++uniq;
                    if (!((t_tl[slot >> 5] >> (slot & 31)) & 1u)) {
                        kl[r] = t_lo[slot];
                        kh[r] = hw;
                        fx128 acc = limb_join(t_l0[slot], t_l1[slot], t_l2[slot]);
                        size_t j = ns;
                        if ((t_tag[slot >> 5] >> (slot & 31)) & 1u) {
                            j = find_source(slo, shi, ns, kl[r], kh[r]);
                            if (j < ns) acc = fx_add(acc, fx_from(dterm[j]));
                            else bad_input = true;
                        }
                        yv[r] = fx_to(acc);
                        em[r] = fabs(yv[r]) > cut;
                        if (j < ns) dacc = fx_add(dacc, fx_from(yv[r] * sv[j]));
                    }
                }
                ne += em[r] ? 1u : 0u;
            }
            {
                const int ntag = min(s_ntag, WR_TAGS);
                em[FR] = false;
                yv[FR] = 0.0;
                kl[FR] = 0;
                kh[FR] = 0;
                if (tid < ntag) {
                    const uint32_t slot = s_tslot[tid], j = s_tj[tid];
                    const fx128 acc = fx_add(limb_join(t_l0[slot], t_l1[slot], t_l2[slot]), fx_from(dterm[j]));
                    yv[FR] = fx_to(acc);
                    em[FR] = fabs(yv[FR]) > cut;
                    kl[FR] = t_lo[slot];
                    kh[FR] = t_hi[slot];
                    dacc = fx_add(dacc, fx_from(yv[FR] * sv[j]));
                    ne += em[FR] ? 1u : 0u;
                }
            }
            uint32_t inc = ne;
            for (int o = 1; o < SFCI_WAVE; o <<= 1) {
                const uint32_t t = __shfl_up(inc, o, SFCI_WAVE);
                if (lane >= o) inc += t;
            }
            if (lane == SFCI_WAVE - 1) s_wtot[tid / SFCI_WAVE] = inc;
            __syncthreads();
            uint32_t pre = inc - ne, tot = 0;
            for (int i = 0; i < WR_THREADS / SFCI_WAVE; ++i) {
                const uint32_t wt = s_wtot[i];
                if (i < tid / SFCI_WAVE) pre += wt;
                tot += wt;
            }
            if (tid == 0) s_obase = tot ? atomicAdd(out_count, (unsigned long long)tot) : 0ull;
            __syncthreads();
            uint64_t pos = s_obase + pre;
#pragma unroll
            for (int r = 0; r <= FR; ++r) {
                if (!em[r]) continue;
                if (pos < out_cap) {
                    olo[pos] = kl[r];
                    ohi[pos] = kh[r];
                    oy[pos] = yv[r];
                } else {
                    overflow = true;
                }
                ++pos;
            }
        }
    };
    // buckets from a global counter (a workgroup's first bucket is its index); the bucket after
    // next is claimed while one is merged, and the next bucket's bounds are loaded meanwhile
    __shared__ unsigned long long s_nb;
    uint64_t b = blockIdx.x;
    if (threadIdx.x == 0) s_nb = gridDim.x + atomicAdd(bucket_next, 1ull);
    __syncthreads();
    uint64_t nb = s_nb;
    uint64_t st = 0, en = 0, nst = 0, nen = 0;
    if (b < n_buckets) {
        st = bstart[b];
        en = bend[b];
    }
    if (nb < n_buckets) {
        nst = bstart[nb];
        nen = bend[nb];
    }
    while (b < n_buckets) {
        unsigned long long nn = 0;
        if (threadIdx.x == 0) nn = gridDim.x + atomicAdd(bucket_next, 1ull);
        if (st < en) merge_bucket(st, en);
        __syncthreads();   // every thread has read s_nb (an empty bucket has no barrier of its own)
        if (threadIdx.x == 0) s_nb = nn;
        __syncthreads();
        const uint64_t nnb = s_nb;
        b = nb;
        st = nst;
        en = nen;
        nb = nnb;
        nst = nen = 0;
        if (nb < n_buckets) {
            nst = bstart[nb];
            nen = bend[nb];
        }
    }
    if (bad_input) atomicOr(flags, FLAG_BAD_INPUT);
    if (overflow) atomicOr(flags, FLAG_OUT_OVERFLOW);
    if (any_full) atomicOr(flags, FLAG_BUCKET_FULL);
    fx_block_reduce_atomic<WR_THREADS>(dacc, diag_acc);
    for (int off = SFCI_WAVE / 2; off > 0; off >>= 1) uniq += __shfl_down(uniq, off, SFCI_WAVE);
    if (lane == 0 && uniq) atomicAdd(unique_count, uniq);
}

// ---- <v|H|v> by pull (calc_energy) ------------------------------------------------------
// Membership table of a key-ordered vector: 2^tb slots, each (32-bit fingerprint = the
// placement hash's top word) << 32 | entry index, or ~0 when empty; open addressing from
// the placement hash's low tb bits.  Keys are distinct, so each entry claims its own slot
// with one compare-and-swap.
__global__ void k_wide_member_build(const uint64_t* __restrict__ k2, size_t n, unsigned long long* __restrict__ table,
                                    int tb) {
    __shared__ uint64_t t_orb[W_ORBS];
    load_orb_words(t_orb);
    __syncthreads();
    const uint64_t mask = ((uint64_t)1 << tb) - 1;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t h = place_hash(t_orb, k2[2 * i], k2[2 * i + 1]);
        const unsigned long long e = ((h >> 32) << 32) | (uint64_t)i;
        uint64_t slot = h & mask;
        while (atomicCAS(&table[slot], WR_EMPTY, e) != WR_EMPTY) slot = (slot + 1) & mask;
    }
}

// index of key (lo, hi) with placement hash h in the vector, or ~0
__device__ __forceinline__ uint64_t member_find(const unsigned long long* __restrict__ table, int tb,
                                                const uint64_t* __restrict__ k2, uint64_t lo, uint64_t hi,
                                                uint64_t h) {
    const uint64_t mask = ((uint64_t)1 << tb) - 1;
    const uint32_t fp = (uint32_t)(h >> 32);
    for (uint64_t slot = h & mask;; slot = (slot + 1) & mask) {
        const unsigned long long e = table[slot];
        if (e == WR_EMPTY) return ~0ull;
        if ((uint32_t)(e >> 32) == fp) {
            const uint64_t i = e & 0xffffffffull;
            if (k2[2 * i] == lo && k2[2 * i + 1] == hi) return i;
        }
    }
}

// sum over i of c_i E_i with E_i = c_i H_ii + sum over connected t in v of H_it c_t: every
// E_i summed exactly and rounded once, c_i E_i summed exactly into acc_out (the 64-bit
// k_energy's rounding; calc_energy, src/functions.cpp:1272-1313)
__global__ __launch_bounds__(W_BLOCK) void k_wide_energy(HamDev H, int n_so, const uint64_t* __restrict__ k2,
                                                         const double* __restrict__ v, size_t n,
                                                         const unsigned long long* __restrict__ table, int tb,
                                                         unsigned long long* acc_out, uint32_t* flags,
                                                         unsigned long long* src_next) {
    __shared__ WSrc s;
    __shared__ uint64_t t_orb[W_ORBS];
    __shared__ unsigned long long s_rlo[W_BLOCK / SFCI_WAVE], s_rhi[W_BLOCK / SFCI_WAVE];
    const int tid = threadIdx.x, lane = tid & (SFCI_WAVE - 1), wv = tid / SFCI_WAVE;
    load_orb_words(t_orb);
    bool bad = false;
    fx128 total = fx_zero();
    // entries from a global counter, as in the generator (the sums are exact: any order)
    __shared__ unsigned long long s_nj;
    unsigned long long nn = 0;
    if (tid == 0) s_nj = gridDim.x + atomicAdd(src_next, 1ull);
    __syncthreads();
    auto next_entry = [&]() -> size_t {
        __syncthreads();
        const size_t jn = (size_t)s_nj;
        __syncthreads();   // every thread has read s_nj
        if (tid == 0) s_nj = nn;
        return jn;
    };
    for (size_t i = blockIdx.x; i < n; i = next_entry()) {
        if (tid == 0) nn = gridDim.x + atomicAdd(src_next, 1ull);
        __syncthreads();
        if (wv == 0) stage_wsrc(H, n_so, k2[2 * i], k2[2 * i + 1], v[i], t_orb, s, flags);
        __syncthreads();
        if (!s.ok) continue;   // block-uniform
        stage_singles(H, s);
        __syncthreads();
        fx128 acc = fx_zero();
        if (wv == 0) {
            bool dok;
            const double e = wide_diag_elem(H, s, dok);
            if (lane == 0) {
                const double t = v[i] * e;
                bad |= !dok || !fx_in_range(t);
                acc = fx_from(t);
            }
        }
        for_connections(H, s, t_orb, [&](uint32_t, uint64_t tl, uint64_t th, uint64_t hx, int p, int q, int x, int y) {
            const uint64_t t = member_find(table, tb, k2, tl, th, hx);
            if (t != ~0ull) {
                const double term = conn_elem(H, s, p, q, x, y) * v[t];
                bad |= !fx_in_range(term);
                acc = fx_add(acc, fx_from(term));
            }
        });
        for (int off = SFCI_WAVE / 2; off > 0; off >>= 1) {
            fx128 o;
            o.lo = __shfl_down(acc.lo, off, SFCI_WAVE);
            o.hi = __shfl_down(acc.hi, off, SFCI_WAVE);
            acc = fx_add(acc, o);
        }
        if (lane == 0) {
            s_rlo[wv] = acc.lo;
            s_rhi[wv] = acc.hi;
        }
        __syncthreads();
        if (tid == 0) {
            fx128 t = fx_zero();
            for (int w = 0; w < W_BLOCK / SFCI_WAVE; ++w) t = fx_add(t, fx128{s_rlo[w], s_rhi[w]});
            const double ci_ei = v[i] * fx_to(t);
            bad |= !fx_in_range(ci_ei);
            total = fx_add(total, fx_from(ci_ei));
        }
    }
    if (tid == 0) fx_atomic_add(acc_out, total);
    if (bad) atomicOr(flags, FLAG_RANGE);
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
    if (ns >= ((uint64_t)1 << 32)) throw std::invalid_argument("more than 2^32 - 1 source determinants");
    double* dterm = c->dterm.get<double>(ns);
    // 2. passes: targets split by the placement hash's top word so that a pass's distinct
    // targets fill 2^16 LDS tables to about WR_PASS_LOAD (no class splits) and its records
    // fit the workspace.  Passes are generated nb at a time into regions of one buffer
    // (26 B per record with its bucket); each region is then partitioned (a gapped copy of
    // part_capacity records, 25 B each) and merged.
    const uint64_t n_conn = H.n_conn;
    const uint64_t total_cand = (uint64_t)ns * n_conn;
    const size_t rec_b = sizeof(RecW) + 2, part_b = sizeof(RecW) + 1;
    auto pass_cap = [&](uint32_t np) -> size_t {
        return np == 1 ? (size_t)total_cand : (size_t)(total_cand / np + total_cand / (8 * np) + 65536);
    };
    auto fits = [&](uint32_t np, int nb) {
        const size_t cp = pass_cap(np);
        return (double)nb * cp * rec_b + (double)part_capacity(cp) * part_b <= (double)c->ws_limit;
    };
    uint32_t n_pass = 1;
    if (c->wide_merge_bits == 0) {
        const double est = (double)total_cand * std::min(1.0, c->unique_ratio * 1.25);
        n_pass = (uint32_t)std::max(1.0, std::ceil(est / (WR_PASS_LOAD * WR_SLOTS * 65536.0)));
    }
    while (!fits(n_pass, 1) && n_pass < (1u << 20)) n_pass = n_pass < 2 ? 2 : n_pass + (n_pass + 1) / 2;
    WStage stg;
    stg.blo = &c->stage_k;
    stg.bhi = &c->lin_k[0];
    stg.by = &c->stage_v;
    stage_reserve(c, stg, std::max<size_t>(out_cap + (out_cap >> 2), (size_t)1 << 16), c->lin_v[0]);
    unsigned long long* d_out = (unsigned long long*)c->d_counter;
    unsigned long long* d_uniq = d_out + 1;
    unsigned long long* d_split = d_out + 3;
    unsigned long long* d_bnext = d_out + 6;   // the merge's bucket counter
    unsigned long long* d_snext = d_out + 7;   // the generator's source counter
    unsigned long long* d_gen = (unsigned long long*)c->src_cnt.get<uint64_t>(W_BATCH);
    const uint32_t fill = c->wide_fill ? (uint32_t)c->wide_fill : (uint32_t)WR_FILL;
    fx128 dsum = fx_zero();
    for (int attempt = 0;; ++attempt) {
        stg.n = 0;
        dsum = fx_zero();
        st.n_candidates = 0;
        st.n_unique = 0;
        st.n_split = 0;
        uint64_t uniq_host = 0;
        HIPCHK(hipMemsetAsync(d_out, 0, 2 * sizeof(uint64_t), c->stream));
        bool overflow = false;
        const size_t cap = pass_cap(n_pass);
        int nb = 1;
        while (n_pass > 1 && nb < W_BATCH && (uint32_t)nb < n_pass && fits(n_pass, nb + 1)) ++nb;
        RecW* A = (RecW*)c->cand_k[0].get<uint8_t>((size_t)nb * cap * sizeof(RecW));
        RecW* B = (RecW*)c->cand_k[1].get<uint8_t>(part_capacity(cap) * sizeof(RecW));
        uint16_t* bkt = c->cand_v[0].get<uint16_t>((size_t)nb * cap);
        // hash bits: the expected distinct targets of a pass at table load 0.6
        int hb = c->wide_merge_bits;
        if (hb == 0) {
            const double est = (double)(total_cand / n_pass) * std::min(1.0, c->unique_ratio * 1.25);
            hb = 9;
            while (hb < 16 && est > 0.6 * WR_SLOTS * (double)(1ull << hb)) ++hb;
        }
        for (uint32_t p0 = 0; p0 < n_pass && !overflow; p0 += (uint32_t)nb) {
            const int nbh = (int)std::min<uint32_t>((uint32_t)nb, n_pass - p0);
            // generate passes [p0, p0 + nbh)
            HIPCHK(hipMemsetAsync(d_gen, 0, W_BATCH * 8, c->stream));
            HIPCHK(hipMemsetAsync(d_snext, 0, 8, c->stream));
            const int gg = (int)std::min<size_t>(ns, 256 * 8);
            int tg = kt_begin(c, K_GENERATE, (double)rec_b * (double)(total_cand / n_pass) * nbh);
            k_wide_generate<MIXED><<<gg, W_BLOCK, 0, c->stream>>>(H, n_so, slo, shi, sv, ns, lambda, coeff, n_pass,
                                                                 p0, nbh, A, bkt, hb, cap, d_gen, dterm,
                                                                 c->d_flags, d_snext);
            kt_end(c, tg);
            HIPCHK(hipGetLastError());
            uint64_t gen_n[W_BATCH] = {};
            if (n_pass == 1) {
                gen_n[0] = total_cand;
            } else {
                HIPCHK(hipMemcpyAsync(gen_n, d_gen, nbh * 8, hipMemcpyDeviceToHost, c->stream));
                HIPCHK(hipStreamSynchronize(c->stream));
            }
            tm.lap(&st.ms_generate);
            const uint32_t f = flags_read_clear(c);
            if (f & FLAG_RANGE) throw HipError("a term left the fixed-point range (|x| >= 2^42)", SFCI_E_RANGE);
            if (f & FLAG_BAD_INPUT)
                throw std::invalid_argument("an input determinant does not have n_elec/2 electrons of each spin "
                                            "inside n_so");
            bool gen_over = (f & FLAG_GEN_OVERFLOW) != 0;
            for (int r = 0; r < nbh; ++r) gen_over |= gen_n[r] > cap;
            if (gen_over) {
                overflow = true;
                break;
            }
            for (int r = 0; r < nbh && !overflow; ++r) {
                const uint64_t n = gen_n[r];
                RecW* Ar = A + (size_t)r * cap;
                st.n_candidates += n;
                if (n == 0) continue;
                // partition by bucket
                const uint64_t nbk = (uint64_t)1 << hb;
                uint64_t* bstart = c->bucket_bounds.get<uint64_t>(2 * nbk);
                uint64_t* bend = bstart + nbk;
                if (!hash_partition_c<RecW>(c, Ar, bkt + (size_t)r * cap, B, n, hb, bstart, bend)) {
                    overflow = true;   // a first-level region overflowed: more passes
                    break;
                }
                tm.lap(&st.ms_sort);
                // merge; a staging list too small for this pass's survivors grows (earlier
                // passes' survivors kept) and the merge runs again
                for (int redo = 0;; ++redo) {
                    const uint64_t before = stg.n;
                    acc_clear(c, 1);
                    HIPCHK(hipMemsetAsync(d_split, 0, 8, c->stream));
                    HIPCHK(hipMemsetAsync(d_bnext, 0, 8, c->stream));
                    int tr = kt_begin(c, K_BUCKET_REDUCE, (double)sizeof(RecW) * n);
                    k_wide_bucket_reduce<<<256, WR_THREADS, 0, c->stream>>>(
                        Ar, bstart, bend, nbk, hb, slo, shi, sv, dterm, ns, c->cut, stg.lo, stg.hi, stg.y, d_out,
                        stg.cap, d_uniq, c->d_acc + 2, c->d_flags, d_split, fill, d_bnext);
                    kt_end(c, tr);
                    HIPCHK(hipGetLastError());
                    uint64_t cnts[3];
                    HIPCHK(hipMemcpyAsync(cnts, d_out, 16, hipMemcpyDeviceToHost, c->stream));
                    HIPCHK(hipMemcpyAsync(cnts + 2, d_split, 8, hipMemcpyDeviceToHost, c->stream));
                    HIPCHK(hipStreamSynchronize(c->stream));
                    const uint32_t fr = flags_read_clear(c);
                    if (fr & FLAG_BAD_INPUT)
                        throw HipError("a source's diagonal target was not found after the merge", SFCI_E_OVERFLOW);
                    if (fr & FLAG_BUCKET_FULL) {
                        overflow = true;   // a bucket needed more than WR_MAX_CLASSES classes: more passes
                        break;
                    }
                    if (cnts[0] <= stg.cap) {
                        stg.n = cnts[0];
                        uniq_host = cnts[1];
                        st.n_split += cnts[2];
                        dsum = fx_add(dsum, acc_read_raw(c, 1));
                        break;
                    }
                    if (redo > 2) throw HipError("merged-output staging overflow after regrow", SFCI_E_OVERFLOW);
                    stage_reserve(c, stg, cnts[0] + (cnts[0] >> 2), c->lin_v[0]);
                    const uint64_t back[2] = {before, uniq_host};
                    HIPCHK(hipMemcpyAsync(d_out, back, 16, hipMemcpyHostToDevice, c->stream));
                    HIPCHK(hipStreamSynchronize(c->stream));
                }
                tm.lap(&st.ms_reduce);
            }
        }
        st.hash_bits = hb;
        st.n_unique = uniq_host;
        if (!overflow) break;
        st.n_fallback += 1;
        if (attempt >= 6) throw HipError("candidate passes kept overflowing the workspace", SFCI_E_OVERFLOW);
        n_pass *= 2;
    }
    st.n_passes = n_pass;
    if (st.n_candidates) c->unique_ratio = std::max(1e-6, (double)st.n_unique / (double)st.n_candidates);
    *diag = fx_to(dsum);
    // 4. output: key order, top out_cap
    const size_t m = stg.n;
    size_t n_out = 0;
    if (m && out_cap) {
        WBufs w2 = wide_bufs(c, m);
        sort128(c, stg.lo, stg.hi, stg.y, m, n_so, w2.a[0], w2.a[1], w2.a[2], w2.a[3], w2.a[4], w2.a[5],
                (double*)w2.a[6]);
        const uint64_t* flo = w2.a[4];
        const uint64_t* fhi = w2.a[5];
        const double* fy = (double*)w2.a[6];
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


// <v|H|v> by pull (calc_energy, src/functions.cpp:1272-1313): v's keys in a hash table,
// then one workgroup per entry i looks up every connection of its key; E_i = c_i H_ii +
// sum H_it c_t summed exactly and rounded once, the c_i E_i summed exactly.
double dev_wide_energy(Ctx* c, const uint64_t* k2, const double* v, size_t n) {
    if (n == 0) return 0.0;
    if (n >= 0xffffffffull) throw std::invalid_argument("more than 2^32 - 2 entries");
    int tb = 10;
    while (((size_t)1 << tb) < 2 * n) ++tb;
    const size_t T = (size_t)1 << tb;
    unsigned long long* table = (unsigned long long*)c->cand_k[0].get<uint64_t>(T);
    (void)flags_read_clear(c);
    k_fill_u64<<<grid_of_w(T), 256, 0, c->stream>>>(table, T, WR_EMPTY);
    k_wide_member_build<<<grid_of_w(n), 256, 0, c->stream>>>(k2, n, table, tb);
    acc_clear(c, 1);
    const int g = (int)std::min<size_t>(n, 256 * 8);
    unsigned long long* d_snext = (unsigned long long*)c->d_counter + 7;   // the pull's entry counter
    HIPCHK(hipMemsetAsync(d_snext, 0, 8, c->stream));
    k_wide_energy<<<g, W_BLOCK, 0, c->stream>>>(c->ham, c->n_so, k2, v, n, table, tb, c->d_acc + 2, c->d_flags,
                                                d_snext);
    HIPCHK(hipGetLastError());
    const double e = acc_read(c, 1);
    const uint32_t f = flags_read_clear(c);
    if (f & FLAG_RANGE) throw HipError("energy: a term left the fixed-point range (|x| >= 2^42)", SFCI_E_RANGE);
    if (f & FLAG_BAD_INPUT)
        throw std::invalid_argument("a determinant does not have n_elec/2 electrons of each spin inside n_so");
    return e;
}

}  // namespace sfci
