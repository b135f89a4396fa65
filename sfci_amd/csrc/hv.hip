// H.v on the GPU: SpMVMtply_bucket (src/functions.cpp:511-804) with 64-bit keys, FP64.
//
// Pipeline for one call:
//   1. input  : top-Nd of the input by |c| when it holds more than Nd entries.
//   2. plan   : target space split into passes by ranges of the target's spin-0
//               string rank (the reference's bucket variable, src/bitstr.cpp:219), so
//               that one pass's candidates fit the workspace.
//   3. per pass
//      count    per-source number of connections landing in the pass range
//      generate one workgroup per source: every connection (diagonal, singles,
//               same-spin and opposite-spin doubles) evaluated once with its full
//               Slater-Condon element; record = (dense target index << 1 | is_diag,
//               H_ij * coeff * c_j), written at deterministic offsets
//      sort     stable LSD radix sort on the pass-local dense index
//      reduce   exact fixed-point sum per target; diag += y_t * c_t for targets in
//               the input; targets with |y| > cut appended to the staging list
//   4. output : top out_cap by (|y| desc, key asc), sorted by key.
#include "internal.h"

namespace sfci {

// spin-orbital integral lookup (m n | p q); zero for spin-forbidden index sets
__device__ __forceinline__ double g_so(const HamDev& H, int m, int n, int p, int q) {
    if (((m ^ n) & 1) | ((p ^ q) & 1)) return 0.0;
    size_t N = (size_t)H.n_orb;
    size_t blk = (size_t)(((m & 1) << 1) | (p & 1));
    return H.g[blk * H.g_stride + (((size_t)(m >> 1) * N + (n >> 1)) * N + (p >> 1)) * N + (q >> 1)];
}

__device__ __forceinline__ double h_so(const HamDev& H, int m, int n) {
    if ((m ^ n) & 1) return 0.0;
    size_t N = (size_t)H.n_orb;
    return H.h[(size_t)(m & 1) * N * N + (size_t)(m >> 1) * N + (n >> 1)];
}

// LDS copies of the small index tables (binomials for colex ranks, pair lists).  Rank
// computations gather from these with per-lane indices; from LDS that is a bank-parallel
// read instead of a scattered L1 request per lane.
constexpr int LDS_BINOM = SFCI_MAX_ORB * 33;
constexpr int LDS_PAIRS = SFCI_MAX_ORB * (SFCI_MAX_ORB - 1) / 2;

__device__ __forceinline__ HamDev stage_tables(const HamDev& H, uint32_t* s_binom, uint16_t* s_op, uint16_t* s_vp) {
    for (int i = threadIdx.x; i < H.n_orb * 33; i += blockDim.x) s_binom[i] = H.binom[i];
    const int nop = H.k * (H.k - 1) / 2, nvp = H.v * (H.v - 1) / 2;
    for (int i = threadIdx.x; i < nop; i += blockDim.x) s_op[i] = H.occ_pairs[i];
    for (int i = threadIdx.x; i < nvp; i += blockDim.x) s_vp[i] = H.vir_pairs[i];
    __syncthreads();
    HamDev L = H;
    L.binom = s_binom;
    L.occ_pairs = s_op;
    L.vir_pairs = s_vp;
    return L;
}

// The generators and the energy pull wait on table reads, lookups and record stores most of
// their cycles: they run at eight wavefronts per SIMD (at most 64 VGPRs; their LDS allows
// eight 256-thread workgroups per CU), which the compiler otherwise trades for a few more
// registers (energy pull of the c2 vector: 301 ms at six wavefronts, 201 ms at eight).  (The
// screened generator's LDS staging allows seven.)
#define GEN_OCCUPANCY __attribute__((amdgpu_waves_per_eu(8, 8)))

constexpr int MAX_SINGLES = (SFCI_MAX_ORB / 2) * (SFCI_MAX_ORB / 2);   // k*v <= 16*16 for 32 orbitals

struct SrcLds {
    uint64_t key;
    double temp;        // coeff * c_j
    double diag_val;    // (E_jj - lambda) * temp
    uint64_t rank[2];
    uint32_t str[2];
    uint8_t occ[2][SFCI_MAX_K];   // spatial orbitals, ascending
    uint8_t vir[2][SFCI_MAX_ORB];
    uint8_t occ_so[2 * SFCI_MAX_K];   // all occupied spin orbitals, ascending
    int n_occ_so;
    uint64_t srank[2][MAX_SINGLES];   // rank of the spin-s string after single excitation e
    uint16_t alist[MAX_SINGLES];      // spin-0 singles whose target row is in the pass range
    // opposite-spin double (spin-0 i -> a with spin-1 j -> b) factors, packed:
    //   bw[e]  spin-1 single e: j*N + b (bits 0-9), min(j,b) (10-14), max(j,b) (15-19),
    //          j (20-24), parity of its single-excitation sign in the source (bit 25)
    //   aw[t]  alist entry t: a*N + i (bits 0-9), i (10-14), a (15-19), sign parity (bit 20)
    //   akey[t] pass-local dense index of the entry's target row, column 0
    uint32_t bw[MAX_SINGLES];
    uint32_t aw[MAX_SINGLES];
    uint64_t akey[MAX_SINGLES];
    int n_alist;
    int row_in;                       // the source's own row is in the pass range
};

// One-time per-source staging.  Thread 0 builds the orbital lists; all threads then
// compute single-excitation ranks and the exact diagonal cooperatively.
__device__ void load_source(const HamDev& H, SrcLds& s, const uint16_t* sdec, uint64_t key, double temp,
                            double lambda, uint64_t row_lo, uint64_t row_hi, uint32_t* flags) {
    __shared__ uint64_t s_dlo[SFCI_BLOCK / SFCI_WAVE], s_dhi[SFCI_BLOCK / SFCI_WAVE];
    __shared__ uint32_t s_acnt[SFCI_BLOCK / SFCI_WAVE];
    const int tid = threadIdx.x, lane = tid & (SFCI_WAVE - 1), wv = tid / SFCI_WAVE;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    if (wv == 0) {
        // one lane per (spin) orbital: occupied / virtual lists and colex ranks by ballots
        const bool occp = lane < 2 * H.n_orb && ((key >> lane) & 1ull);
        const uint64_t bo = __ballot(occp);
        if (occp) s.occ_so[popc64(bo & lt)] = (uint8_t)lane;
        for (int sp = 0; sp < 2; ++sp) {
            const uint32_t st = spin_string(key, sp);
            const bool in = lane < H.n_orb;
            const bool oc = in && ((st >> lane) & 1u);
            const uint64_t b_oc = __ballot(oc), b_vi = __ballot(in && !oc);
            const uint32_t t = (uint32_t)popc64(b_oc & lt);
            if (oc) s.occ[sp][t] = (uint8_t)lane;
            else if (in) s.vir[sp][popc64(b_vi & lt)] = (uint8_t)lane;
            uint64_t r = oc ? H.binom[lane * 33 + t + 1] : 0ull;   // C(pos, index + 1)
            for (int o = SFCI_WAVE / 2; o > 0; o >>= 1) r += __shfl_down(r, o, SFCI_WAVE);
            if (lane == 0) {
                s.str[sp] = st;
                s.rank[sp] = r;
            }
        }
        if (lane == 0) {
            s.key = key;
            s.temp = temp;
            s.n_occ_so = (int)popc64(bo);
            s.row_in = (s.rank[0] >= row_lo && s.rank[0] < row_hi) ? 1 : 0;
            s.n_alist = 0;
        }
    }
    __syncthreads();
    // single-excitation target ranks, both spins
    const int nsa = (int)H.n_sa;
    for (int e = tid; e < 2 * nsa; e += SFCI_BLOCK) {
        int sp = e >= nsa ? 1 : 0, x = e - sp * nsa;
        uint16_t dd = sdec[x];
        const int o = s.occ[sp][dd & 0xff], v = s.vir[sp][dd >> 8];
        uint32_t st = s.str[sp] ^ (1u << o) ^ (1u << v);
        s.srank[sp][x] = colex_rank(st, H.binom);
        if (sp == 1) {
            const uint32_t lo = o < v ? o : v, hi = o < v ? v : o;
            const uint32_t par = sign_between(key, 2 * o + 1, 2 * v + 1) < 0.0 ? 1u : 0u;
            s.bw[x] = ((uint32_t)o * (uint32_t)H.n_orb + (uint32_t)v) | (lo << 10) | (hi << 15) |
                      ((uint32_t)o << 20) | (par << 25);
        }
    }
    // exact diagonal: E_core + sum_occ h_mm + sum_{m != p} 1/2 [(mm|pp) - (mp|pm)], needed
    // only in the pass that holds the source's own row (s.row_in is block-uniform)
    if (s.row_in) {
    FxLimbs lacc = fxl_zero();
    bool dok = true;
    const int n = s.n_occ_so;
    for (int t = tid; t < n * n; t += SFCI_BLOCK) {
        int a = t / n, b = t - a * n;
        int m = s.occ_so[a];
        if (a == b) {
            dok &= fxl_add(lacc, h_so(H, m, m));
        } else {
            int p = s.occ_so[b];
            dok &= fxl_add(lacc, 0.5 * g_so(H, m, m, p, p));
            dok &= fxl_add(lacc, -(0.5 * g_so(H, m, p, p, m)));
        }
    }
    if (!dok) atomicOr(flags, FLAG_RANGE);
    fx128 acc = fxl_join(lacc);
    for (int off = SFCI_WAVE / 2; off > 0; off >>= 1) {
        fx128 o;
        o.lo = __shfl_down(acc.lo, off, SFCI_WAVE);
        o.hi = __shfl_down(acc.hi, off, SFCI_WAVE);
        acc = fx_add(acc, o);
    }
    if ((tid & (SFCI_WAVE - 1)) == 0) {
        s_dlo[tid / SFCI_WAVE] = acc.lo;
        s_dhi[tid / SFCI_WAVE] = acc.hi;
    }
    __syncthreads();
    if (tid == 0) {
        fx128 t = fx_from(H.constant);
        for (int i = 0; i < SFCI_BLOCK / SFCI_WAVE; ++i) t = fx_add(t, fx128{s_dlo[i], s_dhi[i]});
        double dv = (fx_to(t) - lambda) * temp;
        if (!fx_in_range(dv)) atomicOr(flags, FLAG_RANGE);
        s.diag_val = dv;
    }
    }
    // spin-0 singles whose target row is in range, ascending (nsa <= SFCI_BLOCK)
    {
        const bool keep = tid < nsa && s.srank[0][tid] >= row_lo && s.srank[0][tid] < row_hi;
        const uint64_t bk = __ballot(keep);
        if (lane == 0) s_acnt[wv] = (uint32_t)popc64(bk);
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (int i = 0; i < SFCI_BLOCK / SFCI_WAVE; ++i) {
            if (i < wv) pre += s_acnt[i];
            tot += s_acnt[i];
        }
        if (keep) {
            const uint32_t t = pre + (uint32_t)popc64(bk & lt);
            const uint16_t da = sdec[tid];
            const uint32_t i = s.occ[0][da & 0xff], a = s.vir[0][da >> 8];
            const uint32_t par = sign_between(key, 2 * i, 2 * a) < 0.0 ? 1u : 0u;
            s.alist[t] = (uint16_t)tid;
            s.aw[t] = (a * (uint32_t)H.n_orb + i) | (i << 10) | (a << 15) | (par << 20);
            s.akey[t] = s.srank[0][tid] * H.dim - row_lo * H.dim;
        }
        if (tid == 0) s.n_alist = (int)tot;
    }
    __syncthreads();
}

// <D'|H|D>, D' = D with spin orbital n -> m
__device__ __forceinline__ double single_elem(const HamDev& H, const SrcLds& s, int n, int m) {
    FxLimbs acc = fxl_zero();
    bool ok = fxl_add(acc, h_so(H, m, n));
    for (int a = 0; a < s.n_occ_so; ++a) {
        int p = s.occ_so[a];
        if (p == n) continue;
        ok &= fxl_add(acc, g_so(H, m, n, p, p));
        ok &= fxl_add(acc, -g_so(H, m, p, p, n));
    }
    // an integral outside the fixed-point range yields a NaN, which the caller's range
    // check reports
    return ok ? sign_between(s.key, m, n) * fx_to(fxl_join(acc)) : __builtin_nan("");
}

// <D''|H|D>, D'' = D - {n,q} + {x,y}, n < q, x < y (spin orbitals)
__device__ __forceinline__ double double_elem(const HamDev& H, uint64_t d, int n, int q, int x, int y) {
    const uint64_t one = 1;
    uint64_t d1 = d - (one << q) + (one << y);
    double s1 = sign_between(d, q, y) * sign_between(d1, n, x);
    uint64_t d2 = d - (one << q) + (one << x);
    double s2 = sign_between(d, q, x) * sign_between(d2, n, y);
    return s1 * g_so(H, x, n, y, q) + s2 * g_so(H, y, n, x, q);
}

// Opposite-spin double: spin-0 i -> a and spin-1 j -> b (spatial indices).  Same value
// as double_elem(d, n, q, x, y) with n < q, x < y the sorted spin orbitals: exactly one
// of its two terms has a spin-allowed integral, (a i|b j) from the block whose first
// pair is the lower annihilated orbital's spin.
__device__ __forceinline__ double opposite_double(const HamDev& H, uint64_t d, int i, int a, int j, int b) {
    const uint64_t one = 1;
    const int ni = 2 * i, xa = 2 * a, nj = 2 * j + 1, yb = 2 * b + 1;
    const bool n_is_0 = ni < nj;                  // n = ni, q = nj; else n = nj, q = ni
    const int n = n_is_0 ? ni : nj, q = n_is_0 ? nj : ni;
    const bool x_is_0 = xa < yb;
    const int x = x_is_0 ? xa : yb, y = x_is_0 ? yb : xa;
    // the nonzero term is s1*G(x,n,y,q) when x and n share a spin, else s2*G(y,n,x,q)
    const bool first = (n_is_0 == x_is_0);
    double sg;
    if (first) {
        uint64_t d1 = d - (one << q) + (one << y);
        sg = sign_between(d, q, y) * sign_between(d1, n, x);
    } else {
        uint64_t d2 = d - (one << q) + (one << x);
        sg = sign_between(d, q, x) * sign_between(d2, n, y);
    }
    const uint32_t N = (uint32_t)H.n_orb;
    // block (0,1): (a i | b j); block (1,0): (b j | a i); both stored as [a][i][j][b]
    const double g = H.gt[(n_is_0 ? 0u : (uint32_t)H.gt_stride) +
                          (((uint32_t)a * N + (uint32_t)i) * N + (uint32_t)j) * N + (uint32_t)b];
    // the other term is an exact zero; adding it can only change the sign of a zero
    // value, which contributes nothing to any sum
    return sg * g;
}

// Incremental quotient/remainder of i / d for i = start, start + step, ...
struct DivIter {
    uint32_t q, r, dq, dr, d;
    __device__ __forceinline__ DivIter(uint32_t start, uint32_t step, uint32_t div)
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

// ---------------------------------------------------------------------------------
// Per-source count of connections whose target row lies in [row_lo, row_hi).
// Candidate counts of every pass at once, per distinct spin-0 string (every source with
// that string has the same counts): pass p covers target rows [bounds[p], bounds[p + 1]);
// cnt[p * n_row + u] receives the count of a source with string astr[u] in pass p (the
// same weights as k_rowgroup_hist: own row 1 + singles + doubles of spin 1, a spin-0
// single 1 + spin-1 singles, a spin-0 double 1).
constexpr int MAX_PASSES = 256;

__device__ __forceinline__ int pass_of(const uint64_t* bounds, int n_pass, uint64_t r) {
    int lo = 0, hi = n_pass;   // last p with bounds[p] <= r
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (bounds[mid] <= r) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(SFCI_BLOCK) void k_count_passes(HamDev Hg, const uint64_t* __restrict__ astr, size_t n_src,
                                                             const uint64_t* __restrict__ bounds_g, int n_pass,
                                                             uint32_t* __restrict__ cnt) {
    __shared__ uint32_t s_binom[LDS_BINOM];
    __shared__ uint16_t s_op[LDS_PAIRS], s_vp[LDS_PAIRS];
    __shared__ uint64_t bounds[MAX_PASSES + 1];
    __shared__ uint32_t s_cnt[MAX_PASSES];
    __shared__ uint32_t s_str;
    __shared__ uint8_t s_occ[SFCI_MAX_K], s_vir[SFCI_MAX_ORB];
    const HamDev H = stage_tables(Hg, s_binom, s_op, s_vp);
    for (int i = threadIdx.x; i <= n_pass; i += SFCI_BLOCK) bounds[i] = bounds_g[i];
    for (size_t j = blockIdx.x; j < n_src; j += gridDim.x) {
        for (int i = threadIdx.x; i < n_pass; i += SFCI_BLOCK) s_cnt[i] = 0;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t st = (uint32_t)astr[j];
            s_str = st;
            int a = 0, b = 0;
            for (int i = 0; i < H.n_orb; ++i) {
                if (st >> i & 1u) s_occ[a++] = (uint8_t)i;
                else s_vir[b++] = (uint8_t)i;
            }
            atomicAdd(&s_cnt[pass_of(bounds, n_pass, colex_rank(st, H.binom))], (uint32_t)(1 + H.n_sa + H.n_daa));
        }
        __syncthreads();
        const uint64_t n_alpha = H.n_sa + H.n_daa;
        for (uint64_t e = threadIdx.x; e < n_alpha; e += SFCI_BLOCK) {
            uint32_t st;
            uint32_t w;
            if (e < H.n_sa) {
                int ii = (int)(e / H.v), aa = (int)(e - (uint64_t)ii * H.v);
                st = s_str ^ (1u << s_occ[ii]) ^ (1u << s_vir[aa]);
                w = (uint32_t)(1 + H.n_sa);
            } else {
                uint64_t idx = e - H.n_sa;
                uint64_t ij = idx / H.n_vpair, ab = idx - ij * H.n_vpair;
                uint16_t op = H.occ_pairs[ij], vp = H.vir_pairs[ab];
                st = s_str ^ (1u << s_occ[op & 0xff]) ^ (1u << s_occ[op >> 8]) ^ (1u << s_vir[vp & 0xff]) ^
                     (1u << s_vir[vp >> 8]);
                w = 1;
            }
            atomicAdd(&s_cnt[pass_of(bounds, n_pass, colex_rank(st, H.binom))], w);
        }
        __syncthreads();
        for (int p = threadIdx.x; p < n_pass; p += SFCI_BLOCK) cnt[(size_t)p * n_src + j] = s_cnt[p];
        __syncthreads();
    }
}

// Spin orbitals strictly between p and q (the mask sign_between counts in).
__device__ __forceinline__ uint64_t between_mask(int p, int q) {
    const int lo = p < q ? p : q, hi = p < q ? q : p;
    return (hi - lo > 1) ? ((((uint64_t)1 << (hi - lo - 1)) - 1) << (lo + 1)) : 0ull;
}

// Same-spin double tables (SsEnt) of n_str distinct spin-SP strings, one workgroup per
// string: its n_daa doubles in the generator's enumeration order (occupied pair, then
// empty pair), each with the excited string's rank and the element of double_elem split
// into its spin-SP part h and the other-spin parity mask m.  With n_pass > 1 the entries
// are grouped by the target pass of their rank (bounds as in k_count_passes; order inside
// a group unspecified) and pstart[u * (n_pass + 1) + p] receives pass p's first entry.
template <int SP>
__global__ __launch_bounds__(SFCI_BLOCK) void k_ss_table(HamDev Hg, const uint64_t* __restrict__ strs, size_t n_str,
                                                         const uint64_t* __restrict__ bounds_g, int n_pass,
                                                         SsEnt* __restrict__ tab, uint32_t* __restrict__ pstart) {
    __shared__ uint32_t s_binom[LDS_BINOM];
    __shared__ uint16_t s_op[LDS_PAIRS], s_vp[LDS_PAIRS];
    __shared__ uint64_t bounds[MAX_PASSES + 1];
    __shared__ uint32_t s_cnt[MAX_PASSES + 1];
    __shared__ uint8_t s_occ[SFCI_MAX_K], s_vir[SFCI_MAX_ORB];
    const HamDev H = stage_tables(Hg, s_binom, s_op, s_vp);
    const int tid = threadIdx.x;
    const bool grouped = n_pass > 1;
    if (grouped)
        for (int i = tid; i <= n_pass; i += SFCI_BLOCK) bounds[i] = bounds_g[i];
    const uint32_t ndaa = (uint32_t)H.n_daa, nvp = H.n_vpair;
    for (size_t u = blockIdx.x; u < n_str; u += gridDim.x) {
        const uint32_t S = (uint32_t)strs[u];
        __syncthreads();   // the previous string's lists and counters are no longer read
        if (tid == 0) {
            int a = 0, b = 0;
            for (int i = 0; i < H.n_orb; ++i) {
                if (S >> i & 1u) s_occ[a++] = (uint8_t)i;
                else s_vir[b++] = (uint8_t)i;
            }
        }
        if (grouped)
            for (int p = tid; p < n_pass; p += SFCI_BLOCK) s_cnt[p] = 0;
        __syncthreads();
        auto target = [&](const DivIter& it, int& i, int& jo, int& a, int& b) -> uint32_t {
            const uint16_t op = H.occ_pairs[it.q], vp = H.vir_pairs[it.r];
            i = s_occ[op & 0xff];
            jo = s_occ[op >> 8];
            a = s_vir[vp & 0xff];
            b = s_vir[vp >> 8];
            return S ^ (1u << i) ^ (1u << jo) ^ (1u << a) ^ (1u << b);
        };
        if (grouped) {
            DivIter it(tid, SFCI_BLOCK, nvp);
            for (uint32_t f = tid; f < ndaa; f += SFCI_BLOCK, it.next()) {
                int i, jo, a, b;
                const uint64_t r = colex_rank(target(it, i, jo, a, b), H.binom);
                atomicAdd(&s_cnt[pass_of(bounds, n_pass, r)], 1u);
            }
            __syncthreads();
            if (tid == 0) {
                uint32_t run = 0;
                uint32_t* ps = pstart + u * (size_t)(n_pass + 1);
                for (int p = 0; p < n_pass; ++p) {
                    const uint32_t c = s_cnt[p];
                    ps[p] = run;
                    s_cnt[p] = run;   // the pass's cursor
                    run += c;
                }
                ps[n_pass] = run;
            }
            __syncthreads();
        }
        SsEnt* out = tab + u * (size_t)ndaa;
        const uint64_t d = spread_string(S) << SP;
        DivIter it(tid, SFCI_BLOCK, nvp);
        for (uint32_t f = tid; f < ndaa; f += SFCI_BLOCK, it.next()) {
            int i, jo, a, b;
            const uint64_t r = colex_rank(target(it, i, jo, a, b), H.binom);
            const uint32_t slot = grouped ? atomicAdd(&s_cnt[pass_of(bounds, n_pass, r)], 1u) : f;
            const int n = 2 * i + SP, q = 2 * jo + SP, x = 2 * a + SP, y = 2 * b + SP;
            const uint64_t m64 = between_mask(q, y) ^ between_mask(n, x);
            out[slot] = SsEnt{(uint32_t)r, spin_string(m64, 1 - SP), double_elem(H, d, n, q, x, y)};
        }
    }
}

// Per-source index of its string in a distinct-string table built from the string-sorted
// order: sorted position i holds source src_of[i], and starts a new string where fl[i] = 1
// (pos = exclusive scan of fl).
__global__ void k_src_string_index(const double* __restrict__ src_of, const uint32_t* __restrict__ fl,
                                   const uint64_t* __restrict__ pos, size_t n, uint32_t* __restrict__ idx) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        idx[(size_t)src_of[i]] = (uint32_t)(pos[i] + fl[i] - 1);
}

// Histogram of connection counts over target-row groups (row >> grp_shift), for pass
// planning, per distinct spin-0 string astr[j] weighted by its number of sources
// first[j + 1] - first[j].
__global__ __launch_bounds__(SFCI_BLOCK) void k_rowgroup_hist(HamDev Hg, const uint64_t* __restrict__ astr,
                                                              const uint64_t* __restrict__ first, size_t n_src,
                                                              uint64_t row_lo, uint64_t row_hi, int grp_shift, int n_grp,
                                                              unsigned long long* __restrict__ hist) {
    __shared__ uint32_t s_binom[LDS_BINOM];
    __shared__ uint16_t s_op[LDS_PAIRS], s_vp[LDS_PAIRS];
    const HamDev H = stage_tables(Hg, s_binom, s_op, s_vp);
    extern __shared__ unsigned long long s_h[];
    for (int i = threadIdx.x; i < n_grp; i += SFCI_BLOCK) s_h[i] = 0;
    __syncthreads();
    for (size_t j = blockIdx.x; j < n_src; j += gridDim.x) {
        const uint32_t st = (uint32_t)astr[j];
        const unsigned long long m = first[j + 1] - first[j];   // sources with this string
        uint8_t occ[SFCI_MAX_K], vir[SFCI_MAX_ORB];
        int a = 0, b = 0;
        for (int i = 0; i < H.n_orb; ++i) {
            if (st >> i & 1u) occ[a++] = (uint8_t)i;
            else vir[b++] = (uint8_t)i;
        }
        if (threadIdx.x == 0) {
            const uint64_t r = colex_rank(st, H.binom);
            if (r >= row_lo && r < row_hi)
                atomicAdd(&s_h[(r - row_lo) >> grp_shift], m * (unsigned long long)(1 + H.n_sa + H.n_daa));
        }
        const uint64_t n_alpha = H.n_sa + H.n_daa;
        for (uint64_t e = threadIdx.x; e < n_alpha; e += SFCI_BLOCK) {
            uint32_t s2;
            unsigned long long w;
            if (e < H.n_sa) {
                int ii = (int)(e / H.v), aa = (int)(e - (uint64_t)ii * H.v);
                s2 = st ^ (1u << occ[ii]) ^ (1u << vir[aa]);
                w = 1 + H.n_sa;
            } else {
                uint64_t idx = e - H.n_sa;
                uint64_t ij = idx / H.n_vpair, ab = idx - ij * H.n_vpair;
                uint16_t op = H.occ_pairs[ij], vp = H.vir_pairs[ab];
                s2 = st ^ (1u << occ[op & 0xff]) ^ (1u << occ[op >> 8]) ^ (1u << vir[vp & 0xff]) ^ (1u << vir[vp >> 8]);
                w = 1;
            }
            const uint64_t r = colex_rank(s2, H.binom);
            if (r >= row_lo && r < row_hi) atomicAdd(&s_h[(r - row_lo) >> grp_shift], m * w);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n_grp; i += SFCI_BLOCK)
        if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

// String table of the sources' spin-s strings (spin 0: the target rows): keys = strings,
// sorted ascending afterwards with the source index as payload; fl[i] = 1 where a new
// string starts.
__global__ void k_alpha_keys(const uint64_t* __restrict__ sk, size_t n, uint64_t* __restrict__ k, double* __restrict__ v,
                             int spin) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        k[i] = spin_string(sk[i], spin);
        v[i] = (double)i;
    }
}

__global__ void k_row_starts(const uint64_t* __restrict__ k, size_t n, uint32_t* __restrict__ fl) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        fl[i] = (i == 0 || k[i] != k[i - 1]) ? 1u : 0u;
}

// pos = exclusive scan of fl: string u = pos[i] starts at sorted position i
__global__ void k_row_table(const uint64_t* __restrict__ k, const uint32_t* __restrict__ fl,
                            const uint64_t* __restrict__ pos, size_t n, uint64_t n_row, uint64_t* __restrict__ str,
                            uint64_t* __restrict__ first) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (fl[i]) {
            str[pos[i]] = k[i];
            if (first) first[pos[i]] = i;
        }
        if (i == 0 && first) first[n_row] = n;
    }
}

// per-source pass counts from the per-string counts
__global__ void k_expand_counts(const double* __restrict__ src_of, const uint32_t* __restrict__ fl,
                                const uint64_t* __restrict__ pos, size_t n, uint64_t n_row, int n_pass,
                                const uint32_t* __restrict__ cnt_row, uint32_t* __restrict__ cnt) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t j = (size_t)src_of[i];
        const uint64_t u = pos[i] + fl[i] - 1;
        for (int p = 0; p < n_pass; ++p) cnt[(size_t)p * n + j] = cnt_row[(size_t)p * n_row + u];
    }
}

// Bounds-checked record store (the planning count sizes the buffer; a store beyond it is
// dropped and reported).
template <typename R>
__device__ __forceinline__ bool put_rec(R* __restrict__ out, uint64_t cap, uint64_t i, const R& r) {
    if (i >= cap) return false;
    rec_store(out + i, r);
    return true;
}

// Candidate value: the Slater-Condon element times coeff * c_j, in the candidate
// precision (double, or float products for the mixed-precision path).
template <typename T>
__device__ __forceinline__ T cand_val(double hij, double temp);
template <>
__device__ __forceinline__ double cand_val<double>(double hij, double temp) {
    return hij * temp;
}
template <>
__device__ __forceinline__ float cand_val<float>(double hij, double temp) {
    return (float)hij * (float)temp;
}

// Where generated candidates go.  Plain: every record at its planned position.  Screened
// (one pass over all target rows): only records with |value| > eps (diagonal tags always),
// appended to a per-workgroup LDS list that is copied out after each source at an offset
// reserved with one global atomic; a full LDS list spills single records to the global list
// directly.  The screened list's order varies between runs; its content does not.
constexpr uint32_t SCR_LDS = 256;
constexpr uint32_t SCR_FLUSH_AT = 128;   // checkpoints copy the list out once it holds this many
struct ScreenOut {
    RecS* out;
    unsigned long long* count;
    uint64_t cap;
    float eps;
};
template <typename T>
struct PlainSink {
    typename RecOf<T>::type* out;
    uint64_t cap;
    __device__ __forceinline__ bool put(uint64_t idx, uint64_t key, T val, uint32_t, uint64_t) {
        return put_rec(out, cap, idx, make_rec(key, val));
    }
};
// Compact records of a hash-merged pass: the record at its planned position, its hash
// bucket in the 16-bit side array
struct CompactOut {
    void* rec;
    uint16_t* bucket;
    CHash hp;
};
template <typename T>
struct CompactSink {
    typename RecCOf<T>::type* out;
    uint16_t* bucket;
    uint64_t cap;
    CHash hp;
    __device__ __forceinline__ bool put(uint64_t idx, uint64_t key, T val, uint32_t, uint64_t) {
        if (idx >= cap) return false;
        const uint64_t h = ch_hash(hp, key >> 1);
        const uint32_t rem = (uint32_t)(h & ((1ull << hp.rb) - 1));
        rec_store(out + idx, make_crec((rem << 1) | (uint32_t)(key & 1ull), val));
        __builtin_nontemporal_store((uint16_t)(h >> hp.rb), bucket + idx);
        return true;
    }
};
struct ScreenSink {
    ScreenOut so;
    RecS* lds;
    uint32_t* lds_n;
    // one LDS atomic per wavefront call: the first passing lane reserves the wavefront's
    // slots (ballot over the active lanes; divergent callers are fine)
    __device__ __forceinline__ bool put(uint64_t, uint64_t key, float val, uint32_t cls, uint64_t det) {
        const bool pass = (key & 1ull) || fabsf(val) > so.eps;
        const uint64_t bal = __ballot(pass);
        if (!pass) return true;
        const int lane = threadIdx.x & (SFCI_WAVE - 1);
        const int leader = __ffsll((long long)bal) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(lds_n, (uint32_t)popc64(bal));
        base = __shfl(base, leader, SFCI_WAVE);
        const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        const uint32_t i = base + (uint32_t)popc64(bal & lt);
        const RecS r{key, det, val, cls};
        // a full list sends the rest straight to the global list, again one atomic per wavefront
        const bool spill = i >= SCR_LDS;
        const uint64_t sbal = __ballot(spill);
        if (!spill) {
            recs_store<3>(lds + i, r);
            return true;
        }
        const int sl = __ffsll((long long)sbal) - 1;
        unsigned long long gbase = 0;
        if (lane == sl) gbase = atomicAdd(so.count, (unsigned long long)popc64(sbal));
        gbase = __shfl(gbase, sl, SFCI_WAVE);
        const unsigned long long gi = gbase + popc64(sbal & lt);
        if (gi >= so.cap) return false;
        recs_store<1>(so.out + gi, r);
        return true;
    }
    // block-uniform: copy the list out when it holds SCR_FLUSH_AT or more
    __device__ __forceinline__ bool checkpoint(uint64_t* s_base) {
        __syncthreads();
        const bool full = *lds_n >= SCR_FLUSH_AT;
        __syncthreads();
        return full ? flush(s_base) : true;
    }
    // block-uniform: copy the LDS list out and empty it
    __device__ __forceinline__ bool flush(uint64_t* s_base) {
        __syncthreads();
        const uint32_t nl = min(*lds_n, SCR_LDS);
        if (threadIdx.x == 0) *s_base = nl ? atomicAdd(so.count, (unsigned long long)nl) : 0ull;
        __syncthreads();
        bool ok = true;
        for (uint32_t i = threadIdx.x; i < nl; i += blockDim.x) {
            const uint64_t gi = *s_base + i;
            if (gi < so.cap) recs_store<1>(so.out + gi, lds[i]);
            else ok = false;
        }
        __syncthreads();
        if (threadIdx.x == 0) *lds_n = 0;
        __syncthreads();
        return ok;
    }
};
template <typename T, bool SCR, bool CR>
struct SinkOf {
    using type = PlainSink<T>;
};
template <typename T>
struct SinkOf<T, false, true> {
    using type = CompactSink<T>;
};
template <>
struct SinkOf<float, true, false> {
    using type = ScreenSink;
};

constexpr int B2_UNROLL = 4;

// spin-orbital integral (m n | p q) for spin-allowed index sets (the caller knows the spins
// match), 32-bit indexing: four blocks of n_orb^4 <= 2^20 entries
__device__ __forceinline__ double g_so_allowed(const HamDev& H, int m, int n, int p, int q) {
    const uint32_t N = (uint32_t)H.n_orb;
    const uint32_t blk = (uint32_t)(((m & 1) << 1) | (p & 1));
    return H.g[blk * (uint32_t)H.g_stride +
               ((((uint32_t)m >> 1) * N + ((uint32_t)n >> 1)) * N + ((uint32_t)p >> 1)) * N + ((uint32_t)q >> 1)];
}

// Single-excitation elements of one source, n_items of them, one lane each (a source has at
// most k*v <= 256 of each spin): the lane adds h_mn and, over the occupied spin orbitals p != n,
// (m n | p p) and, for p of m's spin, -(m p | p n), in limb form; the same exact sum as
// single_elem.  item(t, n, m, key) gives item t's annihilated/created spin orbitals and
// record key.
template <typename T, typename Item, typename Sink>
__device__ __forceinline__ void emit_singles(const HamDev& H, const SrcLds& s, uint32_t n_items, double temp,
                                             uint64_t pos, Item item, Sink& sink, bool& range_bad, bool& wbad) {
    const int no = s.n_occ_so;
    for (uint32_t t = threadIdx.x; t < n_items; t += SFCI_BLOCK) {
        int n = 0, m = 0;
        uint64_t key = 0;
        item(t, n, m, key);
        FxLimbs acc = fxl_zero();
        bool ok = fxl_add(acc, h_so(H, m, n));
        const int sm = m & 1;
#pragma unroll 2
        for (int a = 0; a < no; ++a) {
            const int p = s.occ_so[a];
            if (p == n) continue;
            const double coul = g_so_allowed(H, m, n, p, p);
            const double exch = (p & 1) == sm ? g_so_allowed(H, m, p, p, n) : 0.0;
            ok &= fxl_add(acc, coul);
            ok &= fxl_add(acc, -exch);
        }
        const double hij = ok ? sign_between(s.key, m, n) * fx_to(fxl_join(acc)) : __builtin_nan("");
        const T val = cand_val<T>(hij, temp);
        range_bad |= !fx_in_range((double)val);
        wbad |= !sink.put(pos + t, key, val, 1u, s.key ^ (1ull << n) ^ (1ull << m));
    }
}

// Candidate generation for the target rows [row_lo, row_hi): one workgroup per source.
//   A0  the diagonal: a tag-only record (value 0); its FP64 term (E_jj - lambda) * temp
//       goes to dterm[j] and is added when the target is merged
//   A1  spin-1 singles, A2  spin-1 doubles          (the source's own row)
//   B1  spin-0 singles whose row is in range
//   B2  their opposite-spin doubles (same rows; string ranks from the LDS tables)
//   C   spin-0 doubles, one row each (rank computed, then range-checked)
// Each part is its own loop, so a wavefront never mixes the O(n_occ) single-excitation
// element with the O(1) double-excitation element.  Records are written at `pos`, which
// starts at j * n_conn (single pass, off == nullptr) or off[j] (per-source exclusive scan
// of the pass counts), so the layout is deterministic.
// With SCR (float only) the candidates go through ScreenSink (screen.eps prefilter) instead
// of the planned positions; it then runs as one pass over all target rows.
template <typename T, bool SCR, bool CR = false>
__device__ __forceinline__ void generate_body(const HamDev& Hg, const uint64_t* __restrict__ sk,
                                              const double* __restrict__ sv, size_t n_src, double lambda, double coeff,
                                              uint64_t row_lo, uint64_t row_hi, const uint64_t* __restrict__ off,
                                              const uint16_t* __restrict__ sdec_g,
                                              typename RecOf<T>::type* __restrict__ out, uint64_t out_cap,
                                              double* __restrict__ dterm, uint32_t* flags, const ScreenOut& screen,
                                              const CompactOut& cout = CompactOut{}, const SsTabs& tabs = SsTabs{},
                                              unsigned long long* src_next = nullptr) {
    __shared__ SrcLds s;
    __shared__ RecS s_scr[SCR ? SCR_LDS : 1];
    __shared__ uint32_t s_scr_n;
    __shared__ uint64_t s_scr_base;
    typename SinkOf<T, SCR, CR>::type sink;
    if constexpr (SCR) {
        sink.so = screen;
        sink.lds = s_scr;
        sink.lds_n = &s_scr_n;
        if (threadIdx.x == 0) s_scr_n = 0;
    } else if constexpr (CR) {
        sink.out = (typename RecCOf<T>::type*)cout.rec;
        sink.bucket = cout.bucket;
        sink.cap = out_cap;
        sink.hp = cout.hp;
    } else {
        sink.out = out;
        sink.cap = out_cap;
    }
    __shared__ uint16_t sdec[MAX_SINGLES];
    __shared__ uint32_t s_ccur;
    __shared__ uint32_t s_binom[LDS_BINOM];
    __shared__ uint16_t s_op[LDS_PAIRS], s_vp[LDS_PAIRS];
    const HamDev H = stage_tables(Hg, s_binom, s_op, s_vp);
    const int tid = threadIdx.x, lane = tid & (SFCI_WAVE - 1), w = tid / SFCI_WAVE;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t base_dense = row_lo * H.dim;
    const uint32_t nsa = (uint32_t)H.n_sa, ndaa = (uint32_t)H.n_daa, nvp = H.n_vpair;
    for (uint32_t x = tid; x < nsa; x += SFCI_BLOCK) sdec[x] = sdec_g[x];
    // classes (highest orbital m of a spin-0 string; its ranks fill [C(m,k), C(m+1,k))) that
    // meet the pass's rows: m1 <= m <= m2
    int m1 = 0, m2 = H.n_orb - 1;
    {
        auto cls_hi = [&](int m) { return m + 1 < H.n_orb ? H.binom[(m + 1) * 33 + H.k] : H.dim; };
        while (m1 < H.n_orb - 1 && cls_hi(m1) <= row_lo) ++m1;
        while (m2 > 0 && H.binom[m2 * 33 + H.k] >= row_hi) --m2;
    }
    // with src_next, sources are taken from a global counter in runs of GEN_CLAIM consecutive
    // sources (a workgroup's first run is its index), so a workgroup on a slower CU takes fewer;
    // the run after next is claimed when a run starts and published when it ends (one atomic
    // per run of 8 when there are at least 64 sources per workgroup: one per source made every
    // workgroup wait on one contended address).  Without, workgroup b takes b, b + grid, ...
    const size_t GEN_CLAIM = n_src >= (size_t)gridDim.x * 64 ? 8 : 1;
    __shared__ unsigned long long s_nj;
    unsigned long long nn = 0;
    if (src_next && threadIdx.x == 0) s_nj = gridDim.x + atomicAdd(src_next, 1ull);
    auto next_source = [&](size_t j) -> size_t {
        if (!src_next) return j + gridDim.x;
        if ((j + 1) % GEN_CLAIM != 0) return j + 1;   // block-uniform
        __syncthreads();
        const size_t jn = (size_t)s_nj * GEN_CLAIM;
        __syncthreads();   // every thread has read s_nj
        if (threadIdx.x == 0) s_nj = nn;
        return jn;
    };
    __syncthreads();
    bool range_bad = false, wbad = false;
    for (size_t j = src_next ? blockIdx.x * GEN_CLAIM : blockIdx.x; j < n_src; j = next_source(j)) {
        if (src_next && threadIdx.x == 0 && j % GEN_CLAIM == 0) nn = gridDim.x + atomicAdd(src_next, 1ull);
        // the string-table indices and this pass's range of the spin-0 table (they depend only
        // on j) are loaded before the source is staged, so their latency overlaps the staging
        uint32_t ta_e0 = 0, ta_e1 = ndaa, ta_u = 0, tb_u = 0;
        if (!SCR && tabs.a != nullptr) {
            ta_u = tabs.a_idx[j];
            if (tabs.n_pass > 1) {
                const uint32_t* ps = tabs.a_start + (size_t)ta_u * (tabs.n_pass + 1) + tabs.pass;
                ta_e0 = ps[0];
                ta_e1 = ps[1];
            }
        }
        if (!SCR && tabs.b != nullptr) tb_u = tabs.b_idx[j];
        load_source(H, s, sdec, sk[j], coeff * sv[j], lambda, row_lo, row_hi, flags);
        const uint64_t d = s.key;
        const double temp = s.temp;
        uint64_t pos = off ? off[j] : (uint64_t)j * H.n_conn;
        if (s.row_in) {
            const uint64_t rowbase = s.rank[0] * H.dim - base_dense;
            // A0
            if (tid == 0) {
                wbad |= !sink.put(pos, (rowbase + s.rank[1]) << 1 | 1ull, tag_val<T>((uint32_t)j), 0u, d);
                dterm[j] = s.diag_val;
            }
            pos += 1;
            // A1
            emit_singles<T>(H, s, nsa, temp, pos,
                            [&](uint32_t f, int& n, int& m, uint64_t& key) {
                                const uint16_t dd = sdec[f];
                                n = 2 * s.occ[1][dd & 0xff] + 1;
                                m = 2 * s.vir[1][dd >> 8] + 1;
                                key = (rowbase + s.srank[1][f]) << 1;
                            },
                            sink, range_bad, wbad);
            pos += nsa;
            // A2
            if (!SCR && tabs.b != nullptr) {
                // from the table of the source's spin-1 string: sign by the spin-0 electrons
                const SsEnt* E = tabs.b + (size_t)tb_u * ndaa;
                const uint32_t s0 = s.str[0];
                const uint32_t dlt = (uint32_t)(pos & 63u);   // wavefront spans start at a multiple of 64 records
                for (uint32_t f1 = tid; f1 < ndaa + dlt; f1 += SFCI_BLOCK) {
                    if (f1 < dlt) continue;
                    const uint32_t f = f1 - dlt;
                    const SsEnt e = E[f];
                    const double hij = (__popc(s0 & e.m) & 1) ? -e.h : e.h;
                    const T val = cand_val<T>(hij, temp);
                    range_bad |= !fx_in_range((double)val);
                    wbad |= !sink.put(pos + f, (rowbase + e.r) << 1, val, 2u, 0ull);
                }
            } else {
            DivIter it(tid, SFCI_BLOCK, nvp);
            for (uint32_t f = tid; f < ndaa; f += SFCI_BLOCK, it.next()) {
                uint16_t op = H.occ_pairs[it.q], vp = H.vir_pairs[it.r];
                int i = s.occ[1][op & 0xff], jo = s.occ[1][op >> 8];
                int a = s.vir[1][vp & 0xff], b = s.vir[1][vp >> 8];
                uint32_t st = s.str[1] ^ (1u << i) ^ (1u << jo) ^ (1u << a) ^ (1u << b);
                T val = cand_val<T>(double_elem(H, d, 2 * i + 1, 2 * jo + 1, 2 * a + 1, 2 * b + 1), temp);
                range_bad |= !fx_in_range((double)val);
                wbad |= !sink.put(pos + f, (rowbase + colex_rank(st, H.binom)) << 1, val, 2u,
                                  d ^ (1ull << (2 * i + 1)) ^ (1ull << (2 * jo + 1)) ^ (1ull << (2 * a + 1)) ^
                                      (1ull << (2 * b + 1)));
            }
            }
            pos += ndaa;
            if constexpr (SCR) wbad |= !sink.checkpoint(&s_scr_base);
        }
        // B1
        const uint32_t na = (uint32_t)s.n_alist;
        emit_singles<T>(H, s, na, temp, pos,
                        [&](uint32_t f, int& n, int& m, uint64_t& key) {
                            const uint32_t aw = s.aw[f];
                            n = 2 * (int)((aw >> 10) & 31u);
                            m = 2 * (int)((aw >> 15) & 31u);
                            key = (s.akey[f] + s.rank[1]) << 1;
                        },
                        sink, range_bad, wbad);
        pos += na;
        if constexpr (SCR) wbad |= !sink.checkpoint(&s_scr_base);
        // B2: <D''|H|D> = (-1)^p (a i|b j), p = parity of the spin-0 single's sign, of the
        // spin-1 single's sign, and of whether 2i and 2a lie between 2j+1 and 2b+1 (moving
        // the spin-0 electron first changes the occupation between the spin-1 pair by
        // exactly those two positions); the same value as opposite_double.  B2_UNROLL
        // integral loads are issued before any is used.
        {
            const uint32_t nB = na * nsa;
            const uint32_t N2 = (uint32_t)H.n_orb * (uint32_t)H.n_orb, gts = (uint32_t)H.gt_stride;
            // plain: item f goes to lane (f + dlt) mod 256 of its round, so that every
            // wavefront's 64 consecutive records start at a record index divisible by 64
            // (whole 128-byte lines of records and of bucket entries: partly written lines
            // cost the memory a read-modify-write)
            const uint32_t dlt = SCR ? 0u : (uint32_t)(pos & 63u);
            const uint32_t nsa_m = nsa > 1 ? 0xFFFFFFFFu / nsa + 1u : 0u;   // f / nsa = umulhi(f, nsa_m), f < 2^16
            // screened: every thread runs every round (block-uniform), so the list can be
            // copied out between rounds; plain: each thread stops at its last item
            for (uint32_t f00 = 0; f00 < nB + dlt; f00 += B2_UNROLL * SFCI_BLOCK) {
                const int f0 = (int)(f00 + tid) - (int)dlt;
                if (!SCR && f0 >= (int)nB) break;
                double g[B2_UNROLL];
                uint32_t qa[B2_UNROLL], qb[B2_UNROLL];
                bool neg[B2_UNROLL];
#pragma unroll
                for (int u = 0; u < B2_UNROLL; ++u) {
                    const int fu = f0 + u * SFCI_BLOCK;
                    g[u] = 0.0;
                    neg[u] = false;
                    qa[u] = qb[u] = 0;
                    if (fu >= 0 && fu < (int)nB) {
                        const uint32_t q = nsa > 1 ? __umulhi((uint32_t)fu, nsa_m) : (uint32_t)fu;
                        qa[u] = q;
                        qb[u] = (uint32_t)fu - q * nsa;
                        const uint32_t aw = s.aw[qa[u]], bw = s.bw[qb[u]];
                        const uint32_t i = (aw >> 10) & 31u, a = (aw >> 15) & 31u;
                        const uint32_t lo = (bw >> 10) & 31u, hi = (bw >> 15) & 31u, j = (bw >> 20) & 31u;
                        const uint32_t p = ((aw >> 20) ^ (bw >> 25) ^ (uint32_t)(i > lo && i <= hi) ^
                                            (uint32_t)(a > lo && a <= hi)) & 1u;
                        neg[u] = p != 0;
                        g[u] = H.gt[(i <= j ? 0u : gts) + (aw & 1023u) * N2 + (bw & 1023u)];
                    }
                }
#pragma unroll
                for (int u = 0; u < B2_UNROLL; ++u) {
                    const int fu = f0 + u * SFCI_BLOCK;
                    if (fu >= 0 && fu < (int)nB) {
                        const uint32_t f = (uint32_t)fu;
                        T val = cand_val<T>(neg[u] ? -g[u] : g[u], temp);
                        range_bad |= !fx_in_range((double)val);
                        uint64_t tdet = 0;
                        if constexpr (SCR) {
                            const uint32_t aw = s.aw[qa[u]], bw = s.bw[qb[u]];
                            const uint32_t i = (aw >> 10) & 31u, a = (aw >> 15) & 31u, jj = (bw >> 20) & 31u;
                            const uint32_t b = ((bw >> 10) & 31u) + ((bw >> 15) & 31u) - jj;
                            tdet = d ^ (1ull << (2 * i)) ^ (1ull << (2 * a)) ^ (1ull << (2 * jj + 1)) ^
                                   (1ull << (2 * b + 1));
                        }
                        wbad |= !sink.put(pos + f, (s.akey[qa[u]] + s.srank[1][qb[u]]) << 1, val, 2u, tdet);
                    }
                }
                if constexpr (SCR) wbad |= !sink.checkpoint(&s_scr_base);
            }
            pos += nB;
        }
        // C (spin-0 doubles in range: each wavefront reserves its slots in the source's range
        // with one LDS atomic; the order of records inside the range is free, every sum
        // downstream being exact)
        if (!SCR && tabs.a != nullptr) {
            // from the table of the source's spin-0 string: this pass's entries, signed by the
            // spin-1 electrons
            const uint32_t u = ta_u, e0 = ta_e0, e1 = ta_e1;
            const SsEnt* E = tabs.a + (size_t)u * ndaa + e0;
            const uint32_t ne = e1 - e0, s1 = s.str[1];
            const uint64_t kb = s.rank[1] - base_dense;   // wraps; the sum with r * dim does not
            const uint32_t dlt = (uint32_t)(pos & 63u);   // wavefront spans start at a multiple of 64 records
            for (uint32_t f1 = tid; f1 < ne + dlt; f1 += SFCI_BLOCK) {
                if (f1 < dlt) continue;
                const uint32_t f = f1 - dlt;
                const SsEnt e = E[f];
                const double hij = (__popc(s1 & e.m) & 1) ? -e.h : e.h;
                const T val = cand_val<T>(hij, temp);
                range_bad |= !fx_in_range((double)val);
                wbad |= !sink.put(pos + f, ((uint64_t)e.r * H.dim + kb) << 1, val, 2u, 0ull);
            }
        } else if (!SCR && off != nullptr) {
            // a row-range pass: only the doubles whose target class meets the pass.  The
            // target's highest orbital is max(b, h) for empty pair (a < b) and h the highest
            // occupied orbital the pair (i, j) leaves; so for each occupied pair, the empty
            // pairs with b in a range, one run of the empty pairs in colex order (by b, then a):
            // s_cb[o] is the run's first colex index, s_cpre the runs' exclusive prefix.
            __shared__ uint16_t s_cb[LDS_PAIRS];
            __shared__ uint32_t s_cpre[LDS_PAIRS + 1];
            const uint32_t nop = H.n_opair;
            const uint32_t vmask = ~s.str[0] & (H.n_orb >= 32 ? ~0u : ((1u << H.n_orb) - 1u));
            auto nv_le = [&](int m) -> uint32_t {   // empty spin-0 orbitals <= m
                if (m < 0) return 0u;
                return (uint32_t)__popc(vmask & (m >= 31 ? ~0u : ((2u << m) - 1u)));
            };
            for (uint32_t o = tid; o < nop; o += SFCI_BLOCK) {
                const uint16_t op = H.occ_pairs[o];
                const int i = op & 0xff, jo = op >> 8;
                int hr = -1;
                for (int t = H.k - 1; t >= 0; --t)
                    if (t != i && t != jo) {
                        hr = s.occ[0][t];
                        break;
                    }
                uint32_t lo = 0, hi = 0;
                if (hr <= m2) {
                    const uint32_t b2 = nv_le(m2), b1 = hr >= m1 ? 0u : nv_le(m1 - 1);
                    lo = b1 * (b1 - (b1 ? 1u : 0u)) / 2;
                    hi = b2 * (b2 - (b2 ? 1u : 0u)) / 2;
                    if (hi < lo) hi = lo;
                }
                s_cb[o] = (uint16_t)lo;
                s_cpre[o] = hi - lo;
            }
            if (tid == 0) s_ccur = 0;
            __syncthreads();
            if (w == 0) {   // exclusive prefix of the run lengths
                uint32_t run = 0;
                for (uint32_t b0 = 0; b0 < nop; b0 += SFCI_WAVE) {
                    const uint32_t o = b0 + lane;
                    const uint32_t len = o < nop ? s_cpre[o] : 0u;
                    uint32_t inc = len;
                    for (int d = 1; d < SFCI_WAVE; d <<= 1) {
                        const uint32_t t = __shfl_up(inc, d, SFCI_WAVE);
                        if (lane >= d) inc += t;
                    }
                    if (o < nop) s_cpre[o] = run + inc - len;
                    run += __shfl(inc, SFCI_WAVE - 1, SFCI_WAVE);
                }
                if (lane == 0) s_cpre[nop] = run;
            }
            __syncthreads();
            const uint32_t total = s_cpre[nop];
            for (uint32_t f0 = 0; f0 < total; f0 += SFCI_BLOCK) {
                const uint32_t f = f0 + tid;
                bool have = false;
                uint64_t key = 0;
                T val = (T)0;
                if (f < total) {
                    uint32_t lo = 0, hi = nop;   // the last occupied pair whose run starts at or before f
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (s_cpre[mid] <= f) lo = mid;
                        else hi = mid;
                    }
                    const uint16_t op = H.occ_pairs[lo];
                    // colex index t -> empty pair (la < lb): lb(lb - 1)/2 <= t < lb(lb + 1)/2
                    const int t = (int)s_cb[lo] + (int)(f - s_cpre[lo]);
                    int lb = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)t)) * 0.5f);
                    if (lb * (lb - 1) / 2 > t) --lb;
                    if ((lb + 1) * lb / 2 <= t) ++lb;
                    const int la = t - lb * (lb - 1) / 2;
                    const int i = s.occ[0][op & 0xff], jo = s.occ[0][op >> 8];
                    const int a = s.vir[0][la], b = s.vir[0][lb];
                    const uint32_t st = s.str[0] ^ (1u << i) ^ (1u << jo) ^ (1u << a) ^ (1u << b);
                    const uint64_t r0 = colex_rank(st, H.binom);
                    have = r0 >= row_lo && r0 < row_hi;
                    if (have) {
                        key = (r0 * H.dim - base_dense + s.rank[1]) << 1;
                        val = cand_val<T>(double_elem(H, d, 2 * i, 2 * jo, 2 * a, 2 * b), temp);
                        range_bad |= !fx_in_range((double)val);
                    }
                }
                const uint64_t bal = __ballot(have);
                if (bal) {
                    const int leader = __ffsll((long long)bal) - 1;
                    uint32_t base = 0;
                    if (lane == leader) base = atomicAdd(&s_ccur, (uint32_t)popc64(bal));
                    base = __shfl(base, leader, SFCI_WAVE);
                    if (have) wbad |= !sink.put(pos + base + (uint32_t)popc64(bal & lt), key, val, 2u, 0ull);
                }
            }
        } else {
            const bool all_in = (off == nullptr);
            if (!all_in) {
                if (tid == 0) s_ccur = 0;
                __syncthreads();
            }
            DivIter it(tid, SFCI_BLOCK, nvp);
            for (uint32_t f0 = 0; f0 < ndaa; f0 += SFCI_BLOCK) {
                uint32_t f = f0 + tid;
                bool have = false;
                uint64_t key = 0, tdet = 0;
                T val = (T)0;
                if (f < ndaa) {
                    uint16_t op = H.occ_pairs[it.q], vp = H.vir_pairs[it.r];
                    int i = s.occ[0][op & 0xff], jo = s.occ[0][op >> 8];
                    int a = s.vir[0][vp & 0xff], b = s.vir[0][vp >> 8];
                    uint32_t st = s.str[0] ^ (1u << i) ^ (1u << jo) ^ (1u << a) ^ (1u << b);
                    if constexpr (SCR)
                        tdet = d ^ (1ull << (2 * i)) ^ (1ull << (2 * jo)) ^ (1ull << (2 * a)) ^ (1ull << (2 * b));
                    // the ranks of k-strings whose highest orbital is m fill [C(m,k), C(m+1,k)):
                    // targets whose class misses the pass range skip the full rank
                    const int top = max(31 - (int)__clz(st), 0);
                    const uint64_t cls_lo = H.binom[top * 33 + H.k];
                    const uint64_t cls_hi = top + 1 < H.n_orb ? H.binom[(top + 1) * 33 + H.k] : H.dim;
                    if (all_in || (cls_hi > row_lo && cls_lo < row_hi)) {
                        const uint64_t r0 = colex_rank(st, H.binom);
                        have = all_in || (r0 >= row_lo && r0 < row_hi);
                        key = r0;
                    }
                    if (have) {
                        const uint64_t r0 = key;
                        key = (r0 * H.dim - base_dense + s.rank[1]) << 1;
                        val = cand_val<T>(double_elem(H, d, 2 * i, 2 * jo, 2 * a, 2 * b), temp);
                        range_bad |= !fx_in_range((double)val);
                    }
                }
                it.next();
                if (all_in) {
                    if (have) wbad |= !sink.put(pos + f, key, val, 2u, tdet);
                    if constexpr (SCR) wbad |= !sink.checkpoint(&s_scr_base);
                    continue;
                }
                const uint64_t bal = __ballot(have);
                if (bal) {
                    const int leader = __ffsll((long long)bal) - 1;
                    uint32_t base = 0;
                    if (lane == leader) base = atomicAdd(&s_ccur, (uint32_t)popc64(bal));
                    base = __shfl(base, leader, SFCI_WAVE);
                    if (have) wbad |= !sink.put(pos + base + (uint32_t)popc64(bal & lt), key, val, 2u, tdet);
                }
            }
        }
        __syncthreads();
        if constexpr (SCR) wbad |= !sink.flush(&s_scr_base);
    }
    if (range_bad) atomicOr(flags, FLAG_RANGE);
    if (wbad) atomicOr(flags, FLAG_GEN_OVERFLOW);
}

template <typename T>
__global__ __launch_bounds__(SFCI_BLOCK) GEN_OCCUPANCY void k_generate(HamDev Hg, const uint64_t* __restrict__ sk,
                                                         const double* __restrict__ sv, size_t n_src, double lambda,
                                                         double coeff, uint64_t row_lo, uint64_t row_hi,
                                                         const uint64_t* __restrict__ off,
                                                         const uint16_t* __restrict__ sdec_g,
                                                         typename RecOf<T>::type* __restrict__ out,
                                                         uint64_t out_cap, double* __restrict__ dterm,
                                                         uint32_t* flags, SsTabs tabs,
                                                         unsigned long long* src_next) {
    generate_body<T, false>(Hg, sk, sv, n_src, lambda, coeff, row_lo, row_hi, off, sdec_g, out, out_cap, dterm, flags,
                            ScreenOut{}, CompactOut{}, tabs, src_next);
}

// the same into compact records and their hash buckets
template <typename T>
__global__ __launch_bounds__(SFCI_BLOCK) GEN_OCCUPANCY void k_generate_compact(HamDev Hg, const uint64_t* __restrict__ sk,
                                                                 const double* __restrict__ sv, size_t n_src,
                                                                 double lambda, double coeff, uint64_t row_lo,
                                                                 uint64_t row_hi, const uint64_t* __restrict__ off,
                                                                 const uint16_t* __restrict__ sdec_g, CompactOut cout,
                                                                 uint64_t out_cap, double* __restrict__ dterm,
                                                                 uint32_t* flags, SsTabs tabs,
                                                                 unsigned long long* src_next) {
    generate_body<T, false, true>(Hg, sk, sv, n_src, lambda, coeff, row_lo, row_hi, off, sdec_g, nullptr, out_cap,
                                  dterm, flags, ScreenOut{}, cout, tabs, src_next);
}

// all target rows, screened (mixed-precision values)
__global__ __launch_bounds__(SFCI_BLOCK) void k_generate_screened(HamDev Hg, const uint64_t* __restrict__ sk,
                                                                  const double* __restrict__ sv, size_t n_src,
                                                                  double lambda, double coeff,
                                                                  const uint16_t* __restrict__ sdec_g,
                                                                  double* __restrict__ dterm, uint32_t* flags,
                                                                  ScreenOut screen, unsigned long long* src_next) {
    generate_body<float, true>(Hg, sk, sv, n_src, lambda, coeff, 0, Hg.dim, nullptr, sdec_g, nullptr, 0, dterm, flags,
                               screen, CompactOut{}, SsTabs{}, src_next);
}

// ---- energy readout: calc_energy (src/functions.cpp:1272-1313) ---------------------
// <v|H|v> = sum_i c_i (c_i H_ii + sum_{t in v, t != i} H_ti c_t), pulled per source: each
// connection of source i is looked up in v (row segment of the target's spin-0 string,
// then a binary search over the sorted spin-1 ranks of that row) and only the members
// are evaluated.  Exact: E_i = round(fx(c_i H_ii) + sum fx(H_ti c_t)), energy =
// round(sum fx(c_i E_i)).  No candidate list is written.
__device__ __forceinline__ uint32_t seg_find(const uint32_t* __restrict__ r1, uint32_t lo, uint32_t hi, uint32_t b) {
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (r1[mid] < b) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(SFCI_BLOCK) GEN_OCCUPANCY void k_energy(HamDev Hg, const uint64_t* __restrict__ sk,
                                                       const double* __restrict__ sv, size_t n_src,
                                                       const uint32_t* __restrict__ row_off,
                                                       const uint32_t* __restrict__ r1s, const double* __restrict__ cs,
                                                       const uint16_t* __restrict__ sdec_g,
                                                       unsigned long long* acc_out, uint32_t* flags,
                                                       unsigned long long* src_next) {
    __shared__ SrcLds s;
    __shared__ uint16_t sdec[MAX_SINGLES];
    __shared__ uint32_t s_binom[LDS_BINOM];
    __shared__ uint16_t s_op[LDS_PAIRS], s_vp[LDS_PAIRS];
    __shared__ uint16_t e_sa[MAX_SINGLES];
    __shared__ uint32_t e_lo[MAX_SINGLES], e_hi[MAX_SINGLES];
    __shared__ uint32_t s_wcnt[SFCI_BLOCK / SFCI_WAVE];
    __shared__ uint64_t s_rlo[SFCI_BLOCK / SFCI_WAVE], s_rhi[SFCI_BLOCK / SFCI_WAVE];
    __shared__ uint32_t s_own[2];
    __shared__ int s_ne;
    const HamDev H = stage_tables(Hg, s_binom, s_op, s_vp);
    const int tid = threadIdx.x, lane = tid & (SFCI_WAVE - 1), w = tid / SFCI_WAVE;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t nsa = (uint32_t)H.n_sa, ndaa = (uint32_t)H.n_daa, nvp = H.n_vpair;
    for (uint32_t x = tid; x < nsa; x += SFCI_BLOCK) sdec[x] = sdec_g[x];
    __syncthreads();
    bool bad = false;
    fx128 total = fx_zero();   // thread 0: sum of fx(c_i E_i) over this block's sources
    // sources from a global counter, as in the generators (the sums are exact: any order)
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
    for (size_t j = blockIdx.x; j < n_src; j = next_source()) {
        if (tid == 0) nn = gridDim.x + atomicAdd(src_next, 1ull);
        load_source(H, s, sdec, sk[j], sv[j], 0.0, 0, H.dim, flags);
        // spin-0 singles whose target row holds members of v (nsa <= SFCI_BLOCK)
        {
            bool have = false;
            uint32_t lo = 0, hi = 0;
            if ((uint32_t)tid < nsa) {
                uint64_t r = s.srank[0][tid];
                lo = row_off[r];
                hi = row_off[r + 1];
                have = hi > lo;
            }
            uint64_t bal = __ballot(have);
            uint32_t my = (uint32_t)popc64(bal & lt);
            if (lane == 0) s_wcnt[w] = (uint32_t)popc64(bal);
            if (tid == 0) {
                s_own[0] = row_off[s.rank[0]];
                s_own[1] = row_off[s.rank[0] + 1];
            }
            __syncthreads();
            uint32_t pre = 0, tot = 0;
            for (int i2 = 0; i2 < SFCI_BLOCK / SFCI_WAVE; ++i2) {
                if (i2 < w) pre += s_wcnt[i2];
                tot += s_wcnt[i2];
            }
            if (have) {
                e_sa[pre + my] = (uint16_t)tid;
                e_lo[pre + my] = lo;
                e_hi[pre + my] = hi;
            }
            if (tid == 0) s_ne = (int)tot;
            __syncthreads();
        }
        const uint64_t d = s.key;
        const uint32_t own_lo = s_own[0], own_hi = s_own[1];
        const uint32_t rb = (uint32_t)s.rank[1];
        fx128 acc = fx_zero();
        if (tid == 0) acc = fx_from(s.diag_val);
        // A1: spin-1 singles (own row)
        for (uint32_t f = tid; f < nsa; f += SFCI_BLOCK) {
            uint32_t b1 = (uint32_t)s.srank[1][f];
            uint32_t x = seg_find(r1s, own_lo, own_hi, b1);
            if (x < own_hi && r1s[x] == b1) {
                uint16_t dd = sdec[f];
                int jo = s.occ[1][dd & 0xff], b = s.vir[1][dd >> 8];
                double t = single_elem(H, s, 2 * jo + 1, 2 * b + 1) * cs[x];
                bad |= !fx_in_range(t);
                acc = fx_add(acc, fx_from(t));
            }
        }
        // A2: spin-1 doubles (own row)
        {
            DivIter it(tid, SFCI_BLOCK, nvp);
            for (uint32_t f = tid; f < ndaa; f += SFCI_BLOCK, it.next()) {
                uint16_t op = H.occ_pairs[it.q], vp = H.vir_pairs[it.r];
                int i = s.occ[1][op & 0xff], jo = s.occ[1][op >> 8];
                int a = s.vir[1][vp & 0xff], b = s.vir[1][vp >> 8];
                uint32_t st = s.str[1] ^ (1u << i) ^ (1u << jo) ^ (1u << a) ^ (1u << b);
                uint32_t b1 = (uint32_t)colex_rank(st, H.binom);
                uint32_t x = seg_find(r1s, own_lo, own_hi, b1);
                if (x < own_hi && r1s[x] == b1) {
                    double t = double_elem(H, d, 2 * i + 1, 2 * jo + 1, 2 * a + 1, 2 * b + 1) * cs[x];
                    bad |= !fx_in_range(t);
                    acc = fx_add(acc, fx_from(t));
                }
            }
        }
        const uint32_t ne = (uint32_t)s_ne;
        // B1: spin-0 singles, spin-1 string unchanged
        for (uint32_t f = tid; f < ne; f += SFCI_BLOCK) {
            uint32_t x = seg_find(r1s, e_lo[f], e_hi[f], rb);
            if (x < e_hi[f] && r1s[x] == rb) {
                uint16_t da = sdec[e_sa[f]];
                int i = s.occ[0][da & 0xff], a = s.vir[0][da >> 8];
                double t = single_elem(H, s, 2 * i, 2 * a) * cs[x];
                bad |= !fx_in_range(t);
                acc = fx_add(acc, fx_from(t));
            }
        }
        // B2: opposite-spin doubles over the populated spin-0 rows
        {
            const uint32_t nB = ne * nsa;
            DivIter it(tid, SFCI_BLOCK, nsa ? nsa : 1);
            for (uint32_t f = tid; f < nB; f += SFCI_BLOCK, it.next()) {
                uint32_t e = it.q, sb = it.r;
                uint32_t b1 = (uint32_t)s.srank[1][sb];
                uint32_t x = seg_find(r1s, e_lo[e], e_hi[e], b1);
                if (x < e_hi[e] && r1s[x] == b1) {
                    uint16_t da = sdec[e_sa[e]], db = sdec[sb];
                    int i = s.occ[0][da & 0xff], a = s.vir[0][da >> 8];
                    int jo = s.occ[1][db & 0xff], b = s.vir[1][db >> 8];
                    double t = opposite_double(H, d, i, a, jo, b) * cs[x];
                    bad |= !fx_in_range(t);
                    acc = fx_add(acc, fx_from(t));
                }
            }
        }
        // C: spin-0 doubles, spin-1 string unchanged
        {
            DivIter it(tid, SFCI_BLOCK, nvp);
            for (uint32_t f = tid; f < ndaa; f += SFCI_BLOCK, it.next()) {
                uint16_t op = H.occ_pairs[it.q], vp = H.vir_pairs[it.r];
                int i = s.occ[0][op & 0xff], jo = s.occ[0][op >> 8];
                int a = s.vir[0][vp & 0xff], b = s.vir[0][vp >> 8];
                uint32_t st = s.str[0] ^ (1u << i) ^ (1u << jo) ^ (1u << a) ^ (1u << b);
                uint64_t r0 = colex_rank(st, H.binom);
                uint32_t lo = row_off[r0], hi = row_off[r0 + 1];
                if (hi > lo) {
                    uint32_t x = seg_find(r1s, lo, hi, rb);
                    if (x < hi && r1s[x] == rb) {
                        double t = double_elem(H, d, 2 * i, 2 * jo, 2 * a, 2 * b) * cs[x];
                        bad |= !fx_in_range(t);
                        acc = fx_add(acc, fx_from(t));
                    }
                }
            }
        }
        // E_i = round(sum), then c_i E_i into the block total
        for (int off = SFCI_WAVE / 2; off > 0; off >>= 1) {
            fx128 o;
            o.lo = __shfl_down(acc.lo, off, SFCI_WAVE);
            o.hi = __shfl_down(acc.hi, off, SFCI_WAVE);
            acc = fx_add(acc, o);
        }
        if (lane == 0) {
            s_rlo[w] = acc.lo;
            s_rhi[w] = acc.hi;
        }
        __syncthreads();
        if (tid == 0) {
            fx128 t = fx_zero();
            for (int i = 0; i < SFCI_BLOCK / SFCI_WAVE; ++i) t = fx_add(t, fx128{s_rlo[i], s_rhi[i]});
            double ci_ei = s.temp * fx_to(t);
            bad |= !fx_in_range(ci_ei);
            total = fx_add(total, fx_from(ci_ei));
        }
        __syncthreads();
    }
    if (tid == 0) fx_atomic_add(acc_out, total);
    if (bad) atomicOr(flags, FLAG_RANGE);
}

// membership index of v: spin-1 ranks sorted inside each spin-0 row, row r spanning
// [row_off[r], row_off[r + 1])
__global__ void k_dense_split(HamDev H, const uint64_t* __restrict__ dets, const double* __restrict__ c, size_t n,
                              uint64_t* __restrict__ dk, double* __restrict__ dv) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint64_t r0 = colex_rank(spin_string(dets[i], 0), H.binom);
        uint64_t r1 = colex_rank(spin_string(dets[i], 1), H.binom);
        dk[i] = r0 * H.dim + r1;
        dv[i] = c[i];
    }
}

__global__ void k_row_index(const uint64_t* __restrict__ dk, size_t n, uint64_t dim, uint32_t* __restrict__ row_off,
                            uint32_t* __restrict__ r1s) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r <= dim; r += stride) {
        // first entry with row >= r
        size_t lo = 0, hi = n;
        const uint64_t key = r * dim;
        while (lo < hi) {
            size_t mid = (lo + hi) >> 1;
            if (dk[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        row_off[r] = (uint32_t)lo;
    }
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) r1s[i] = (uint32_t)(dk[i] % dim);
}

template <typename R, typename T>
__global__ void k_rec_split(const R* __restrict__ rec, size_t n, uint64_t* __restrict__ k, T* __restrict__ v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const R x = rec[i];
        k[i] = rec_key(x);
        v[i] = (rec_key(x) & 1ull) ? (T)0 : rec_val(x);   // a tag record's value is its source index
    }
}

// Finish one merged target: for a target that is an input determinant (tag), add the
// FP64 diagonal term of its source and accumulate y_t * c_t into diag; round the exact
// sum; decide the cut.  Kept targets are staged under their dense index (converted to
// determinant keys once, after every pass); the key is rebuilt here only for the search of
// a tagged target.
__device__ __forceinline__ double finish_target(const HamDev& H, uint64_t dense, fx128 acc, uint32_t tag,
                                                const uint64_t* __restrict__ ik, const double* __restrict__ iv,
                                                const double* __restrict__ dterm, size_t n_in, double cut, bool& emit,
                                                fx128& dacc, bool& bad_input) {
    size_t lo = 0;
    bool found = false;
    if (tag) {
        uint64_t r0 = dense / H.dim, r1 = dense - r0 * H.dim;
        const uint64_t det = spread_string(colex_unrank(r0, H.k, H.n_orb, H.binom)) |
                             (spread_string(colex_unrank(r1, H.k, H.n_orb, H.binom)) << 1);
        size_t hi = n_in;
        while (lo < hi) {
            size_t mid = (lo + hi) >> 1;
            if (ik[mid] < det) lo = mid + 1;
            else hi = mid;
        }
        found = lo < n_in && ik[lo] == det;
        if (found) acc = fx_add(acc, fx_from(dterm[lo]));
        else bad_input = true;
    }
    double y = fx_to(acc);
    emit = fabs(y) > cut;
    if (found) dacc = fx_add(dacc, fx_from(y * iv[lo]));
    return y;
}

// Exact per-target merge of sorted candidates (segmented reduction).
//
// Block b owns the targets whose first candidate lies in tile b (RED_TILE elements); it
// reads on past the tile end to finish its last target.  The owned range is consumed in
// chunks of RED_CHUNK: every thread sums RED_IPT consecutive candidates sequentially,
// then a block-wide segmented scan over threads carries partial sums of targets that
// straddle threads (and chunks).  Sums are exact 128-bit fixed point, so the grouping
// does not affect the result.
constexpr int RED_IPT = 4;
constexpr int RED_CHUNK = SFCI_BLOCK * RED_IPT;
constexpr int RED_TILE = 4 * RED_CHUNK;

struct SegVal {
    fx128 v;
    uint32_t f;   // a target starts inside the range
    uint32_t t;   // a diagonal (input) contribution seen
};

__device__ __forceinline__ SegVal seg_combine(const SegVal& a, const SegVal& b) {
    SegVal r;
    r.f = a.f | b.f;
    if (b.f) {
        r.v = b.v;
        r.t = b.t;
    } else {
        r.v = fx_add(a.v, b.v);
        r.t = a.t | b.t;
    }
    return r;
}

__device__ __forceinline__ SegVal seg_shfl_up(const SegVal& x, int off) {
    SegVal r;
    r.v.lo = __shfl_up(x.v.lo, off, SFCI_WAVE);
    r.v.hi = __shfl_up(x.v.hi, off, SFCI_WAVE);
    r.f = __shfl_up(x.f, off, SFCI_WAVE);
    r.t = __shfl_up(x.t, off, SFCI_WAVE);
    return r;
}

template <typename T>
__global__ __launch_bounds__(SFCI_BLOCK) void k_reduce(HamDev H, const uint64_t* __restrict__ ck,
                                                       const T* __restrict__ cv, size_t n, uint64_t base_dense,
                                                       const uint64_t* __restrict__ ik, const double* __restrict__ iv,
                                                       const double* __restrict__ dterm,
                                                       size_t n_in, double cut, uint64_t* __restrict__ sk_out,
                                                       double* __restrict__ sv_out, unsigned long long* out_count,
                                                       size_t out_cap, unsigned long long* unique_count,
                                                       unsigned long long* diag_acc, uint32_t* flags) {
    __shared__ unsigned long long s_start, s_end;
    __shared__ SegVal s_wave[SFCI_BLOCK / SFCI_WAVE];
    __shared__ SegVal s_incl[SFCI_BLOCK];
    const int tid = threadIdx.x, lane = tid & (SFCI_WAVE - 1), w = tid / SFCI_WAVE;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    fx128 dacc = fx_zero();
    unsigned long long uniq = 0;
    bool bad_input = false, overflow = false;
    for (size_t tile = blockIdx.x; tile * RED_TILE < n; tile += gridDim.x) {
        const size_t tb = tile * RED_TILE, te = (tb + RED_TILE < n) ? tb + RED_TILE : n;
        if (tid == 0) {
            s_start = te;
            s_end = n;
        }
        __syncthreads();
        for (size_t i = tb + tid; i < te; i += SFCI_BLOCK) {
            if (i == 0 || (ck[i] >> 1) != (ck[i - 1] >> 1)) {
                atomicMin(&s_start, (unsigned long long)i);
                break;
            }
        }
        for (size_t i0 = te; i0 < n; i0 += SFCI_BLOCK) {
            size_t i = i0 + tid;
            if (i < n && (ck[i] >> 1) != (ck[i - 1] >> 1)) atomicMin(&s_end, (unsigned long long)i);
            __syncthreads();
            bool done = s_end < n;
            __syncthreads();
            if (done) break;
        }
        __syncthreads();
        const size_t start = s_start, end = s_end;
        __syncthreads();
        if (start >= te) continue;   // the whole tile continues a target owned by an earlier block
        SegVal carry{fx_zero(), 0u, 0u};
        for (size_t c0 = start; c0 < end; c0 += RED_CHUNK) {
            const size_t b = c0 + (size_t)tid * RED_IPT;
            fx128 first = fx_zero(), tailv = fx_zero();
            uint32_t ftag = 0, ttag = 0;
            bool has_head = false, first_ends = false;
            uint64_t first_key = 0;
            uint64_t ekey[RED_IPT];
            fx128 eval[RED_IPT];
            uint32_t etag[RED_IPT];
            int ne = 0;
#pragma unroll
            for (int r = 0; r < RED_IPT; ++r) {
                size_t i = b + r;
                if (i >= end) break;
                uint64_t raw = ck[i];
                uint64_t key = raw >> 1;
                bool head = (i == start) || (key != (ck[i - 1] >> 1));
                bool ends = (i + 1 == end) || ((ck[i + 1] >> 1) != key);
                fx128 x = fx_from((double)cv[i]);
                uint32_t tg = (uint32_t)(raw & 1ull);
                if (head) {
                    has_head = true;
                    tailv = x;
                    ttag = tg;
                } else if (!has_head) {
                    first = fx_add(first, x);
                    ftag |= tg;
                } else {
                    tailv = fx_add(tailv, x);
                    ttag |= tg;
                }
                if (ends) {
                    if (has_head) {
                        ekey[ne] = key;
                        eval[ne] = tailv;
                        etag[ne] = ttag;
                        ++ne;
                    } else {
                        first_ends = true;
                        first_key = key;
                    }
                }
            }
            // segmented inclusive scan over threads, seeded with the chunk carry
            SegVal x{has_head ? tailv : first, has_head ? 1u : 0u, has_head ? ttag : ftag};
            for (int off = 1; off < SFCI_WAVE; off <<= 1) {
                SegVal o = seg_shfl_up(x, off);
                if (lane >= off) x = seg_combine(o, x);
            }
            if (lane == SFCI_WAVE - 1) s_wave[w] = x;
            __syncthreads();
            SegVal pre = carry;
            for (int j = 0; j < w; ++j) pre = seg_combine(pre, s_wave[j]);
            s_incl[tid] = seg_combine(pre, x);
            __syncthreads();
            const SegVal cin = tid > 0 ? s_incl[tid - 1] : carry;
            carry = s_incl[SFCI_BLOCK - 1];
            __syncthreads();
            if (first_ends) {
                // the target that entered this thread ends here: prepend the carried part
                for (int q = ne; q > 0; --q) {
                    ekey[q] = ekey[q - 1];
                    eval[q] = eval[q - 1];
                    etag[q] = etag[q - 1];
                }
                ekey[0] = first_key;
                eval[0] = fx_add(cin.v, first);
                etag[0] = cin.t | ftag;
                ++ne;
            }
            // finish the targets: round, diagonal contribution, cut, append
#pragma unroll
            for (int r = 0; r < RED_IPT; ++r) {
                bool emit = false;
                uint64_t det = 0;
                double y = 0.0;
                if (r < ne) {
                    ++uniq;
                    det = ekey[r] + base_dense;   // staged as the dense index
                    y = finish_target(H, det, eval[r], etag[r], ik, iv, dterm, n_in, cut, emit, dacc, bad_input);
                }
                uint64_t bal = __ballot(emit);
                if (bal) {
                    unsigned long long base = 0;
                    if (lane == 0) base = atomicAdd(out_count, (unsigned long long)popc64(bal));
                    base = __shfl(base, 0, SFCI_WAVE);
                    if (emit) {
                        uint64_t pos = base + (uint64_t)popc64(bal & lt);
                        if (pos < out_cap) {
                            sk_out[pos] = det;
                            sv_out[pos] = y;
                        } else {
                            overflow = true;
                        }
                    }
                }
            }
        }
    }
    if (bad_input) atomicOr(flags, FLAG_BAD_INPUT);
    if (overflow) atomicOr(flags, FLAG_OUT_OVERFLOW);
    fx_block_reduce_atomic<SFCI_BLOCK>(dacc, diag_acc);
    for (int off = SFCI_WAVE / 2; off > 0; off >>= 1) uniq += __shfl_down(uniq, off, SFCI_WAVE);
    if (lane == 0 && uniq) atomicAdd(unique_count, uniq);
}

// ---------------------------------------------------------------------------------
// Hash-bucket merge.  After radix_partition_hash the candidates are grouped by the top
// hb bits of target_hash(target); a bucket holds about a few thousand distinct targets,
// which one workgroup merges in an LDS open-addressing table with exact 128-bit
// accumulators (64-bit LDS atomics with carry).  When a bucket holds more distinct targets
// than the table takes at BR_FILL occupancy, the workgroup splits it into hash classes
// (8 more hash bits below the slot bits, taken modulo a power of two) and merges one class
// at a time, rereading the bucket's records for each; every finished class is a complete
// set of targets, so classes split further when they overflow in turn.  Only a class that
// still overflows at 256 classes sets FLAG_BUCKET_FULL (the pass then takes the full sort).
constexpr int BR_THREADS = 1024;
constexpr int BR_SLOTS = 4608;                 // full-key tables: 4608 x 32 B = 144 KB of LDS per workgroup
constexpr int BR_SLOTS_C = 5504;               // compact-key tables: 5504 x 28 B = 150.5 KB (with the tag
                                               // lists and class stacks about 158 of the CU's 160 KB)
template <bool CR>
constexpr int br_slots() {
    return CR ? BR_SLOTS_C : BR_SLOTS;
}
// records per thread loaded together: as many as the 128 registers of a 1024-thread
// workgroup hold without spilling
template <typename T, bool CR>
constexpr int br_batch() {
    return (!CR && sizeof(T) == 8) ? 6 : 4;
}
constexpr int BR_TAGS = 512;                   // tagged targets per bucket finished from a list
constexpr int BR_FILL = BR_SLOTS * 4 / 5;      // occupied slots at which a table counts as full
constexpr int BR_FILL_C = BR_SLOTS_C * 4 / 5;
constexpr int BR_STACK = 256;                  // pending classes of one bucket
constexpr uint32_t BR_MAX_CLASSES = 4096;      // classes of one bucket beyond which the pass takes the full sort
static_assert(BR_TAGS <= BR_THREADS, "the tag list is finished one entry per thread");
static_assert(BR_SLOTS_C % 4 == 0, "compact-key tables are probed in aligned groups of four slots");
constexpr unsigned long long BR_EMPTY = ~0ull;

// Bucket b of a hash-partitioned candidate list (ascending in target_hash(key >> 1, hb))
// spans [lower_bound(b), lower_bound(b + 1)): one binary search per bound instead of a
// pass over the keys.
template <typename R>
__device__ __forceinline__ size_t hash_lower_bound(const R* __restrict__ rec, size_t n, int hb, uint64_t b) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = (lo + hi) >> 1;
        if (target_hash(rec_key(rec[mid]) >> 1, hb) < b) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <typename R>
__global__ void k_bucket_bounds(const R* __restrict__ ck, size_t n, int hb, uint64_t n_buckets,
                                uint64_t* __restrict__ bstart, uint64_t* __restrict__ bend) {
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < n_buckets;
         b += (uint64_t)gridDim.x * blockDim.x) {
        bstart[b] = hash_lower_bound(ck, n, hb, b);
        bend[b] = b + 1 < n_buckets ? hash_lower_bound(ck, n, hb, b + 1) : n;
    }
}

// CR: compact records (the bucket index and the record key give the target's hash, hp
// inverts it); otherwise records carry the pass-local dense target.
template <typename T, bool CR>
__global__ __launch_bounds__(BR_THREADS) void k_bucket_reduce(
    HamDev H, const void* __restrict__ rec_in, CHash hp, const uint64_t* __restrict__ bstart,
    const uint64_t* __restrict__ bend, uint64_t n_buckets, int hb, uint64_t base_dense,
    const uint64_t* __restrict__ ik, const double* __restrict__ iv, const double* __restrict__ dterm, size_t n_in,
    double cut,
    uint64_t* __restrict__ sk_out, double* __restrict__ sv_out, unsigned long long* out_count, size_t out_cap,
    unsigned long long* unique_count, unsigned long long* diag_acc, uint32_t* flags,
    unsigned long long* split_count, uint32_t fill_max, unsigned long long* bucket_next) {
    // table keys: compact records carry a 31-bit key, stored as one 32-bit word (four slots
    // per 128-bit LDS read); full-key records a 64-bit dense index
    using KT = typename std::conditional<CR, uint32_t, unsigned long long>::type;
    constexpr KT EMPTYK = CR ? (KT)0xFFFFFFFFu : (KT)BR_EMPTY;
    constexpr int SLOTS = br_slots<CR>();
    __shared__ __align__(16) KT t_key[SLOTS];
    __shared__ unsigned long long t_l0[SLOTS], t_l1[SLOTS], t_l2[SLOTS];
    __shared__ uint32_t t_tag[(SLOTS + 31) / 32];   // tagged, finished by search
    __shared__ uint32_t t_tl[(SLOTS + 31) / 32];    // tagged, finished from the tag list
    __shared__ uint32_t s_tslot[BR_TAGS], s_tj[BR_TAGS];
    __shared__ int s_full, s_ntag, s_sp, s_bad;
    __shared__ uint32_t s_fill;
    __shared__ uint32_t s_nread;
    __shared__ uint32_t s_cm[BR_STACK], s_cq[BR_STACK];   // pending classes: (count, index)
    __shared__ uint32_t s_wtot[BR_THREADS / SFCI_WAVE];
    __shared__ unsigned long long s_obase;
    using R = typename std::conditional<CR, typename RecCOf<T>::type, typename RecOf<T>::type>::type;
    constexpr int BRB = br_batch<T, CR>();
    const R* __restrict__ rec = (const R*)rec_in;
    const int tid = threadIdx.x, lane = tid & (SFCI_WAVE - 1);
    const int slot_shift = 48 - hb;             // 16 hash bits below the bucket bits
    const int cls_shift = 32 - hb;              // the class value: 16 hash bits below the slot bits
    if (tid == 0) s_bad = 0;
    uint64_t b = blockIdx.x;                    // the bucket being merged
    // table key -> dense target index
    auto to_dense = [&](unsigned long long key) -> uint64_t {
        if constexpr (CR) return base_dense + ch_unhash(hp, (b << hp.rb) | key);
        else return base_dense + key;
    };
    fx128 dacc = fx_zero();
    unsigned long long uniq = 0;
    bool bad_input = false, overflow = false, any_full = false;
    // one bucket: merge its records [st, en) in the LDS table and finish its targets
    // one bucket: merge its records [st, en) class by class in the LDS table and finish
    // their targets; the class stack starts with the whole bucket (modulus 1, class 0)
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
            uint32_t n_ins = 0;     // records of this class this thread inserted
            // BRB records per thread per round, all loaded before any is inserted; full
            // rounds load without bounds tests, so that their loads are in flight together
            constexpr uint64_t STEP = (uint64_t)BRB * BR_THREADS;
            auto load_full = [&](uint64_t i0, uint64_t* k_, T* v_) {
#pragma unroll
                for (int r = 0; r < BRB; ++r) {
                    const auto x = rec[i0 + (uint64_t)r * BR_THREADS + tid];
                    k_[r] = rec_key(x);
                    v_[r] = rec_val(x);
                }
            };
            auto load_round = [&](uint64_t i0, uint64_t* k_, T* v_) {
#pragma unroll
                for (int r = 0; r < BRB; ++r) {
                    const uint64_t i = i0 + (uint64_t)r * BR_THREADS + tid;
                    k_[r] = ~0ull;
                    v_[r] = (T)0;
                    if (i < en) {
                        const auto x = rec[i];
                        k_[r] = rec_key(x);
                        v_[r] = rec_val(x);
                    }
                }
            };
            typedef uint32_t k4_t __attribute__((ext_vector_type(4)));
            // the aligned four-slot group of compact keys holding slot; read through an LDS-space
            // pointer (ds_read_b128): a generic pointer would make it a flat load that waits for
            // all outstanding global loads
            auto read_group = [&](uint32_t slot) -> k4_t {
                return *(volatile const __attribute__((address_space(3))) k4_t*)(
                    (const __attribute__((address_space(3))) uint32_t*)(t_key + (slot & ~3u)));
            };
            // the table counts as full once fill_max slots are taken (the fill counter's adds
            // return nothing; it is read at round boundaries and every 16 probes) or a probe
            // saw every group
            auto table_full = [&]() -> bool {
                return __hip_atomic_load(&s_full, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) ||
                       __hip_atomic_load(&s_fill, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= fill_max;
            };
            auto insert_round = [&](const uint64_t* k_, const T* v_) {
                // compact keys: the first group of every record of the round is read before any
                // is probed, so that the reads' latencies overlap
                k4_t kf[BRB];
                uint32_t sf[BRB];
                if constexpr (CR) {
#pragma unroll
                    for (int r = 0; r < BRB; ++r) {
                        const uint32_t t = (uint32_t)(k_[r] >> 1) * 0x9E3779B1u;
                        sf[r] = (uint32_t)(((t >> 16) * (uint32_t)SLOTS) >> 16);
                        if (k_[r] != ~0ull) kf[r] = read_group(sf[r]);
                    }
                }
#pragma unroll
                for (int r = 0; r < BRB; ++r) {
                    if (k_[r] == ~0ull) continue;
                    const uint64_t raw = k_[r];
                    const KT key = (KT)(raw >> 1);
                    uint32_t s16, c16;   // 16 hash bits for the slot, the next 16 for the class
                    if constexpr (CR) {
                        // the key's rb bits mixed into 32 (a bijection mod 2^32): the top 16
                        // bits for the slot; the low 16 for the class, distinct for every key
                        // of a bucket that has at most 2^16 keys
                        const uint32_t t = (uint32_t)key * 0x9E3779B1u;
                        s16 = t >> 16;
                        c16 = t & 0xffffu;
                    } else {
                        const uint64_t hf = target_hash_full(key);
                        s16 = (uint32_t)(hf >> slot_shift) & 0xffffu;
                        c16 = (uint32_t)(hf >> cls_shift) & 0xffffu;
                    }
                    // class cq of cm: the cm-th part of the 16-bit class value's range
                    if (cm > 1 && (uint32_t)((c16 * cm) >> 16) != cq) continue;
                    ++n_ins;
                    uint32_t slot = (uint32_t)((s16 * (uint32_t)SLOTS) >> 16);
                    // most records meet a key already in the table (a target collects one
                    // record per connected source): a plain read finds it, and only an empty
                    // slot costs a compare-and-swap.  A slot's key never changes once set.
                    int probe = 0;
                    if constexpr (CR) {
                        // linear probing four slots at a time: one 128-bit read of the aligned
                        // group holding `slot` (the first one read above); the first of (our
                        // key, an empty slot) at or after `slot` in probe order decides.  A group
                        // read earlier may be stale: a key inserted since is met by the CAS.
                        for (;;) {
                            if ((probe & 15) == 15 && table_full()) {
                                probe = SLOTS;   // the class is being abandoned: give up the search
                                break;
                            }
                            if (probe * 4 > SLOTS + 4) {   // every group seen
                                probe = SLOTS;
                                break;
                            }
                            const uint32_t g = slot & ~3u, q0 = slot & 3u;
                            const k4_t kk = probe == 0 ? kf[r] : read_group(slot);
                            ++probe;
                            uint32_t hit = 4, emp = 4;
#pragma unroll
                            for (int q = 3; q >= 0; --q) {
                                if ((uint32_t)q >= q0) {
                                    if (kk[q] == key) hit = (uint32_t)q;
                                    if (kk[q] == EMPTYK) emp = (uint32_t)q;
                                }
                            }
                            if (hit < emp) {
                                slot = g + hit;
                                break;
                            }
                            if (emp < 4) {
                                slot = g + emp;
                                const KT prev = atomicCAS(&t_key[slot], EMPTYK, key);
                                if (prev == EMPTYK) {
                                    atomicAdd(&s_fill, 1u);
                                    break;
                                }
                                if (prev == key) break;
                                slot = (slot + 1 == (uint32_t)SLOTS) ? 0u : slot + 1;   // taken by another key
                                continue;
                            }
                            slot = (g + 4 >= (uint32_t)SLOTS) ? 0u : g + 4;
                        }
                    } else {
                    for (; probe < SLOTS; ++probe) {
                        if ((probe & 63) == 63 && table_full()) {
                            probe = SLOTS;   // the class is being abandoned: give up the search
                            break;
                        }
                        const KT cur = __hip_atomic_load(&t_key[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (cur == key) break;
                        if (cur == EMPTYK) {
                            const KT prev = atomicCAS(&t_key[slot], EMPTYK, key);
                            if (prev == EMPTYK) {
                                atomicAdd(&s_fill, 1u);
                                break;
                            }
                            if (prev == key) break;
                        }
                        slot = (slot + 1 == (uint32_t)SLOTS) ? 0u : slot + 1;
                    }
                    }
                    if (probe == SLOTS) {
                        s_full = 1;
                        continue;
                    }
                    if (raw & 1ull) {
                        // the target's own source: (slot, source index) into the tag list, or,
                        // when that is full, a tag bit for the search in finish_target
                        const int t = atomicAdd(&s_ntag, 1);
                        if (t < BR_TAGS) {
                            s_tslot[t] = slot;
                            s_tj[t] = tag_index(v_[r]);
                            atomicOr(&t_tl[slot >> 5], 1u << (slot & 31));
                        } else {
                            atomicOr(&t_tag[slot >> 5], 1u << (slot & 31));
                        }
                        continue;
                    }
                    // l0 and l1 are added unconditionally (a zero add costs less than the
                    // branch); l2 is zero for every |term| < 1 and is skipped then
                    unsigned long long l0, l1, l2;
                    limb_split((double)v_[r], l0, l1, l2);
                    atomicAdd(&t_l0[slot], l0);
                    atomicAdd(&t_l1[slot], l1);
                    if (l2) atomicAdd(&t_l2[slot], l2);
                }
            };
            // the first round's loads are issued before the table is cleared
            uint64_t ka[BRB];
            T va[BRB];
            if (st + STEP <= en) load_full(st, ka, va);
            else load_round(st, ka, va);
            for (int i = tid; i < SLOTS; i += BR_THREADS) {
                t_key[i] = EMPTYK;
                t_l0[i] = 0;
                t_l1[i] = 0;
                t_l2[i] = 0;
            }
            for (int i = tid; i < (SLOTS + 31) / 32; i += BR_THREADS) {
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
            insert_round(ka, va);
            uint64_t i0 = st + STEP;
            // a full table ends the class early: each thread stops at its next round
            for (; i0 + STEP <= en && !table_full(); i0 += STEP) {
                load_full(i0, ka, va);
                insert_round(ka, va);
            }
            if (i0 < en && !table_full()) {
                load_round(i0, ka, va);
                insert_round(ka, va);
                i0 = en;
            }
            __syncthreads();
            const bool full = s_full != 0 || s_fill >= fill_max;
            if (full) {
                // split this class into f subclasses (class cq of cm -> classes cq*f .. cq*f+f-1
                // of cm*f, a refinement of its range).  Its targets are estimated from the
                // table's fill after n_read of its records: with records drawn over D keys,
                // fill = D (1 - exp(-n_read / D)), solved for D by bisection.
                uint32_t r = n_ins;
                for (int o = SFCI_WAVE / 2; o > 0; o >>= 1) r += __shfl_down(r, o, SFCI_WAVE);
                if (lane == 0 && r) atomicAdd(&s_nread, r);
                __syncthreads();
                if (tid == 0) {
                    uint32_t f = split_count_for(s_fill, s_nread, SLOTS * 3 / 5, BR_STACK);
                    if (f < 2) f = 2;
                    if (s_sp + (int)f > BR_STACK) f = (uint32_t)(BR_STACK - s_sp);
                    if (f < 2 || cm * f > BR_MAX_CLASSES) {
                        s_bad = 1;      // too many classes (or no stack room): the pass takes the full sort
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
            // finish: every slot's target, and the tag list's, is rounded and cut; the kept
            // ones (staged under their dense index) are written after one block-wide scan of
            // the per-thread counts, with one global reservation per class
            constexpr int FR = (SLOTS + BR_THREADS - 1) / BR_THREADS;
            bool em[FR + 1];
            double yv[FR + 1];
            uint64_t dk[FR + 1];
            uint32_t ne = 0;
#pragma unroll
            for (int r = 0; r < FR; ++r) {
                const int slot = r * BR_THREADS + tid;
                em[r] = false;
                yv[r] = 0.0;
                dk[r] = 0;
                const KT key = slot < SLOTS ? t_key[slot] : EMPTYK;
                if (key != EMPTYK) {
                    ++uniq;
                    if (!((t_tl[slot >> 5] >> (slot & 31)) & 1u)) {
                        const uint32_t tag = (t_tag[slot >> 5] >> (slot & 31)) & 1u;
                        dk[r] = to_dense(key);
                        yv[r] = finish_target(H, dk[r], limb_join(t_l0[slot], t_l1[slot], t_l2[slot]), tag, ik,
                                              iv, dterm, n_in, cut, em[r], dacc, bad_input);
                    }
                }
                ne += em[r] ? 1u : 0u;
            }
            // targets that are input determinants, from the tag list (BR_TAGS <= BR_THREADS):
            // the same finish as finish_target with the source index known, no search
            {
                const int ntag = min(s_ntag, BR_TAGS);
                em[FR] = false;
                yv[FR] = 0.0;
                dk[FR] = 0;
                if (tid < ntag) {
                    const uint32_t slot = s_tslot[tid], j = s_tj[tid];
                    const fx128 acc = fx_add(limb_join(t_l0[slot], t_l1[slot], t_l2[slot]), fx_from(dterm[j]));
                    yv[FR] = fx_to(acc);
                    em[FR] = fabs(yv[FR]) > cut;
                    dk[FR] = to_dense(t_key[slot]);
                    dacc = fx_add(dacc, fx_from(yv[FR] * iv[j]));
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
            for (int i = 0; i < BR_THREADS / SFCI_WAVE; ++i) {
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
                    sk_out[pos] = dk[r];
                    sv_out[pos] = yv[r];
                } else {
                    overflow = true;
                }
                ++pos;
            }
        }
    };
    // buckets are taken from a global counter (a workgroup's first bucket is its index), so a
    // workgroup slowed by class splits takes fewer; the bucket after next is claimed while this
    // one is merged, and the next bucket's bounds are loaded meanwhile
    __shared__ unsigned long long s_nb;
    if (tid == 0) s_nb = gridDim.x + atomicAdd(bucket_next, 1ull);
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
        if (tid == 0) nn = gridDim.x + atomicAdd(bucket_next, 1ull);
        if (st < en) merge_bucket(st, en);
        __syncthreads();   // every thread has read s_nb (an empty bucket has no barrier of its own)
        if (tid == 0) s_nb = nn;
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
    fx_block_reduce_atomic<BR_THREADS>(dacc, diag_acc);
    for (int off = SFCI_WAVE / 2; off > 0; off >>= 1) uniq += __shfl_down(uniq, off, SFCI_WAVE);
    if (lane == 0 && uniq) atomicAdd(unique_count, uniq);
}

// Staged outputs carry dense indices (spin-0 rank * dim + spin-1 rank); rewrite them as
// determinant keys, with the binomial table in LDS.
__global__ __launch_bounds__(SFCI_BLOCK) void k_dense_to_det(HamDev Hg, uint64_t* __restrict__ k, size_t n) {
    __shared__ uint32_t s_binom[LDS_BINOM];
    for (int i = threadIdx.x; i < Hg.n_orb * 33; i += blockDim.x) s_binom[i] = Hg.binom[i];
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t dense = k[i];
        const uint64_t r0 = dense / Hg.dim, r1 = dense - r0 * Hg.dim;
        k[i] = spread_string(colex_unrank(r0, Hg.k, Hg.n_orb, s_binom)) |
               (spread_string(colex_unrank(r1, Hg.k, Hg.n_orb, s_binom)) << 1);
    }
}

// ---------------------------------------------------------------------------------

static int grid_for(Ctx* c, size_t work, int per_cu = 8) {
    int cus = 256;
    (void)c;
    size_t g = std::min<size_t>(work, (size_t)cus * per_cu);
    return (int)std::max<size_t>(g, 1);
}


template <typename T>
static size_t dev_hv_t(Ctx* c, const uint64_t* ik, const double* iv, size_t n_in, size_t nd, double lambda,
                       double coeff, size_t out_cap, uint64_t* ok, double* ov, double* diag, HvAll* all) {
    const HamDev& H = c->ham;
    auto& st = c->st;
    kt_resolve(c);   // laps left pending by an earlier call that threw
    st = {};
    PhaseTimer total(c);
    *diag = 0.0;
    if (n_in == 0) return 0;
    // 1. input: top-Nd
    const uint64_t* sk = ik;
    const double* sv = iv;
    size_t ns = n_in;
    if (n_in > nd) {
        uint64_t* tk = c->src_k.get<uint64_t>(nd);
        double* tv = c->src_v.get<double>(nd);
        ns = select_top(c, ik, iv, n_in, nd, tk, tv);
        sk = tk;
        sv = tv;
    }
    st.n_src = ns;
    double* dterm = c->dterm.get<double>(ns);
    // 2. plan passes over target rows
    const uint64_t n_rows = H.dim;
    const size_t rec = sizeof(uint64_t) + sizeof(double);
    // two record buffers per pass and one digit byte per first-level slot
    const size_t pass_cap = std::max<size_t>(c->ws_limit / (2 * rec + 1), (size_t)1 << 20);
    // keep the distinct targets of a pass within 2^16 LDS-table buckets at load c->pass_load
    // (0.9 at most: more passes cost more generation than an occasional class split).  A row
    // segment's distinct targets are estimated from its candidates and the distinct/candidate
    // ratio the previous call saw in that row range (x 1.25); a bucket that overflows anyway is
    // merged in hash classes, each rereading its records, and the load aimed at adapts to how
    // many buckets split in the previous call
    const double dcap = c->merge_mode == 0 ? c->pass_load * BR_SLOTS_C * 65536.0 : INFINITY;
    auto ratio_at = [&](uint64_t row) -> double {
        double r = c->unique_ratio;
        if (!c->rr_lo.empty()) {
            const size_t i = (size_t)(std::upper_bound(c->rr_lo.begin(), c->rr_lo.end(), row) - c->rr_lo.begin());
            r = c->rr_ratio[i ? i - 1 : 0];
        }
        return std::min(1.0, r * 1.25);
    };
    uint64_t total_cand = (uint64_t)ns * H.n_conn;
    // distinct spin-0 strings of the sources (planning work depends only on them)
    uint64_t n_row_src = 0;
    uint64_t *row_str = nullptr, *row_first = nullptr, *row_pos = nullptr;
    uint32_t* row_fl = nullptr;
    double* row_src = nullptr;
    auto row_table = [&]() {
        if (row_str) return;
        uint64_t* ka = c->plan_k[0].get<uint64_t>(ns);
        double* va = c->plan_v[0].get<double>(ns);
        uint64_t* kb = c->plan_k[1].get<uint64_t>(ns);
        double* vb = c->plan_v[1].get<double>(ns);
        const int g = grid_for(c, (ns + SFCI_BLOCK - 1) / SFCI_BLOCK, 16);
        k_alpha_keys<<<g, SFCI_BLOCK, 0, c->stream>>>(sk, ns, ka, va, 0);
        radix_sort_pairs(c, &ka, &va, kb, vb, ns, H.n_orb);
        row_fl = c->plan_fl.get<uint32_t>(ns);
        row_pos = c->plan_pos.get<uint64_t>(ns + 1);
        k_row_starts<<<g, SFCI_BLOCK, 0, c->stream>>>(ka, ns, row_fl);
        exclusive_scan_u32(c, row_fl, ns, row_pos, row_pos + ns);
        HIPCHK(hipMemcpyAsync(&n_row_src, row_pos + ns, 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        row_str = c->plan_str.get<uint64_t>(n_row_src);
        row_first = c->plan_first.get<uint64_t>(n_row_src + 1);
        k_row_table<<<g, SFCI_BLOCK, 0, c->stream>>>(ka, row_fl, row_pos, ns, n_row_src, row_str, row_first);
        HIPCHK(hipGetLastError());
        row_src = va;
    };
    std::vector<std::pair<uint64_t, uint64_t>> passes;
    std::vector<uint64_t> pass_count;   // candidates of each pass
    std::vector<double> pass_dist;      // distinct targets estimated for each pass
    if (total_cand <= pass_cap && (double)total_cand * std::min(1.0, c->unique_ratio * 1.25) <= dcap) {
        passes.push_back({0, n_rows});
        pass_count.push_back(total_cand);
        pass_dist.push_back((double)total_cand * std::min(1.0, c->unique_ratio * 1.25));
    } else {
        // candidates per row segment: a histogram over at most 4096 row groups, refined
        // inside any group that alone exceeds the pass capacity, down to single rows
        struct Seg {
            uint64_t lo, hi, count;
        };
        auto hist_of = [&](uint64_t lo, uint64_t hi) {
            int grp_shift = 0;
            while (((hi - lo - 1) >> grp_shift) >= 4096) ++grp_shift;
            const int n_grp = (int)((hi - lo - 1) >> grp_shift) + 1;
            unsigned long long* hist = c->rowhist.get<unsigned long long>(n_grp);
            HIPCHK(hipMemsetAsync(hist, 0, n_grp * sizeof(unsigned long long), c->stream));
            row_table();
            int tr = kt_begin(c, K_ROWGROUP, 16.0 * n_row_src);
            k_rowgroup_hist<<<grid_for(c, n_row_src), SFCI_BLOCK, n_grp * sizeof(unsigned long long), c->stream>>>(
                H, row_str, row_first, n_row_src, lo, hi, grp_shift, n_grp, hist);
            kt_end(c, tr);
            std::vector<unsigned long long> hh(n_grp);
            HIPCHK(hipMemcpyAsync(hh.data(), hist, n_grp * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                  c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            std::vector<Seg> out(n_grp);
            for (int g = 0; g < n_grp; ++g)
                out[g] = Seg{lo + ((uint64_t)g << grp_shift), std::min(hi, lo + ((uint64_t)(g + 1) << grp_shift)),
                             hh[g]};
            return out;
        };
        // a segment too large for one pass by either count
        auto too_big = [&](const Seg& sg) {
            return sg.count > pass_cap || (double)sg.count * ratio_at(sg.lo) > dcap;
        };
        std::vector<Seg> segs = hist_of(0, n_rows);
        for (int round = 0; round < 8; ++round) {
            bool refined = false;
            std::vector<Seg> next;
            for (const Seg& sg : segs) {
                if (too_big(sg) && sg.hi - sg.lo > 1) {
                    std::vector<Seg> sub = hist_of(sg.lo, sg.hi);
                    next.insert(next.end(), sub.begin(), sub.end());
                    refined = true;
                } else {
                    next.push_back(sg);
                }
            }
            segs.swap(next);
            if (!refined) break;
        }
        uint64_t lo = 0, acc = 0;
        double acc_d = 0.0;   // distinct targets estimated for the pass being filled
        for (const Seg& sg : segs) {
            if (sg.count > pass_cap)
                throw HipError("one target row exceeds the pass capacity; raise the workspace limit", -4);
            const double d = (double)sg.count * ratio_at(sg.lo);
            if (acc > 0 && (acc + sg.count > pass_cap || acc_d + d > dcap)) {
                passes.push_back({lo, sg.lo});
                pass_count.push_back(acc);
                pass_dist.push_back(acc_d);
                lo = sg.lo;
                acc = 0;
                acc_d = 0.0;
            }
            acc += sg.count;
            acc_d += d;
        }
        passes.push_back({lo, n_rows});
        pass_count.push_back(acc);
        pass_dist.push_back(acc_d);
    }
    st.n_passes = passes.size();
    uint32_t* pcnt = nullptr;
    uint64_t* dbounds = nullptr;
    std::vector<uint64_t> bounds;
    if (passes.size() > 1) {
        if (passes.size() > (size_t)MAX_PASSES) throw HipError("more target passes than supported; raise the workspace", -4);
        bounds.resize(passes.size() + 1);
        for (size_t p = 0; p < passes.size(); ++p) bounds[p] = passes[p].first;
        bounds[passes.size()] = n_rows;
        dbounds = c->part_meta[1].get<uint64_t>(bounds.size());
        HIPCHK(hipMemcpyAsync(dbounds, bounds.data(), bounds.size() * 8, hipMemcpyHostToDevice, c->stream));
        pcnt = c->src_cnt.get<uint32_t>(passes.size() * ns);
        row_table();
        uint32_t* pcnt_row = c->plan_cnt.get<uint32_t>(passes.size() * n_row_src);
        int tcn = kt_begin(c, K_ROWGROUP, 8.0 * n_row_src + 4.0 * passes.size() * (n_row_src + ns));
        k_count_passes<<<grid_for(c, n_row_src), SFCI_BLOCK, 0, c->stream>>>(H, row_str, n_row_src, dbounds,
                                                                             (int)passes.size(), pcnt_row);
        k_expand_counts<<<grid_for(c, (ns + SFCI_BLOCK - 1) / SFCI_BLOCK, 16), SFCI_BLOCK, 0, c->stream>>>(
            row_src, row_fl, row_pos, ns, n_row_src, (int)passes.size(), pcnt_row, pcnt);
        kt_end(c, tcn);
        HIPCHK(hipGetLastError());
    }
    // same-spin double tables: each distinct source string's doubles (excited rank, element)
    // computed once, so the generator's spin-0 doubles (class C) and spin-1 doubles (A2)
    // stream them instead of enumerating and evaluating every double of every source in
    // every pass.  Automatic when that work is large and the tables fit in a quarter of the
    // workspace limit.
    SsTabs tabs{};
    tabs.n_pass = (int)passes.size();
    st.n_str0 = st.n_str1 = 0;
    if (H.n_daa > 0 && (c->gen_tables == 2 || (c->gen_tables == 0 && (uint64_t)ns * H.n_daa >= (1ull << 22)))) {
        const size_t budget = c->gen_tables == 2 ? SIZE_MAX : c->ws_limit / 4;
        const size_t ent = sizeof(SsEnt) * (size_t)H.n_daa;
        const int g = grid_for(c, (ns + SFCI_BLOCK - 1) / SFCI_BLOCK, 16);
        row_table();
        uint64_t n1 = 0;
        uint64_t* str1 = nullptr;
        double* src1 = nullptr;
        uint32_t *fl1 = nullptr;
        uint64_t* pos1 = nullptr;
        {
            // distinct spin-1 strings, as row_table does for spin 0
            uint64_t* ka = c->bstr_k[0].get<uint64_t>(ns);
            double* va = c->bstr_v[0].get<double>(ns);
            uint64_t* kb = c->bstr_k[1].get<uint64_t>(ns);
            double* vb = c->bstr_v[1].get<double>(ns);
            k_alpha_keys<<<g, SFCI_BLOCK, 0, c->stream>>>(sk, ns, ka, va, 1);
            radix_sort_pairs(c, &ka, &va, kb, vb, ns, H.n_orb);
            fl1 = c->bstr_fl.get<uint32_t>(ns);
            pos1 = c->bstr_pos.get<uint64_t>(ns + 1);
            k_row_starts<<<g, SFCI_BLOCK, 0, c->stream>>>(ka, ns, fl1);
            exclusive_scan_u32(c, fl1, ns, pos1, pos1 + ns);
            HIPCHK(hipMemcpyAsync(&n1, pos1 + ns, 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            str1 = c->bstr_str.get<uint64_t>(n1);
            k_row_table<<<g, SFCI_BLOCK, 0, c->stream>>>(ka, fl1, pos1, ns, n1, str1, nullptr);
            src1 = va;
        }
        const int np = (int)passes.size();
        const size_t bytes0 = n_row_src * ent + (np > 1 ? n_row_src * (size_t)(np + 1) * sizeof(uint32_t) : 0);
        const size_t bytes1 = n1 * ent;
        int tk = kt_begin(c, K_SS_TABLE, (double)(bytes0 + bytes1) + 8.0 * ns);
        if (bytes0 <= budget) {
            SsEnt* t0 = c->ss_tab[0].get<SsEnt>(n_row_src * H.n_daa);
            uint32_t* ps = np > 1 ? c->ss_start.get<uint32_t>(n_row_src * (size_t)(np + 1)) : nullptr;
            uint32_t* ix = c->ss_idx[0].get<uint32_t>(ns);
            k_src_string_index<<<g, SFCI_BLOCK, 0, c->stream>>>(row_src, row_fl, row_pos, ns, ix);
            k_ss_table<0><<<grid_for(c, n_row_src), SFCI_BLOCK, 0, c->stream>>>(H, row_str, n_row_src, dbounds, np, t0,
                                                                                  ps);
            tabs.a = t0;
            tabs.a_start = ps;
            tabs.a_idx = ix;
            st.n_str0 = n_row_src;
        }
        if (bytes1 <= (budget > bytes0 ? budget - bytes0 : 0) || (tabs.a == nullptr && bytes1 <= budget)) {
            SsEnt* t1 = c->ss_tab[1].get<SsEnt>(n1 * H.n_daa);
            uint32_t* ix = c->ss_idx[1].get<uint32_t>(ns);
            k_src_string_index<<<g, SFCI_BLOCK, 0, c->stream>>>(src1, fl1, pos1, ns, ix);
            k_ss_table<1><<<grid_for(c, n1), SFCI_BLOCK, 0, c->stream>>>(H, str1, n1, nullptr, 1, t1, nullptr);
            tabs.b = t1;
            tabs.b_idx = ix;
            st.n_str1 = n1;
        }
        kt_end(c, tk);
        HIPCHK(hipGetLastError());
    }
    if (passes.size() > 1) HIPCHK(hipStreamSynchronize(c->stream));   // bounds live on the host stack
    // staging for merged output
    size_t stage_cap = std::max<size_t>(out_cap + (out_cap >> 2), (size_t)1 << 16);
    if (all) {
        // every distinct target is kept: size from the last observed distinct/candidate ratio
        double est = (double)total_cand * std::min(1.0, c->unique_ratio * 1.15) + 65536.0;
        stage_cap = std::max<size_t>(stage_cap, (size_t)std::min<double>(est, (double)total_cand + 1.0));
    }
    if (c->stage_k.bytes / sizeof(uint64_t) > stage_cap) stage_cap = c->stage_k.bytes / sizeof(uint64_t);
    uint64_t* stk = c->stage_k.get<uint64_t>(stage_cap);
    double* stv = c->stage_v.get<double>(stage_cap);
    // no output wanted (energy): nothing passes the cut; all targets wanted: everything does
    const double cut_eff = all ? -1.0 : (out_cap ? c->cut : INFINITY);
    fx128 diag_total = fx_zero();
    uint64_t uniq_host = 0;
    unsigned long long* d_out_count = (unsigned long long*)c->d_counter;
    unsigned long long* d_uniq = d_out_count + 1;
    HIPCHK(hipMemsetAsync(c->d_counter, 0, 2 * sizeof(uint64_t), c->stream));
    uint64_t out_count_host = 0;
    uint64_t buckets_merged = 0;   // hash buckets of the passes merged through LDS tables
    std::vector<uint64_t> rr_lo;   // this call's distinct/candidate ratio per pass, for the next
    std::vector<double> rr_ratio;
    for (size_t pi = 0; pi < passes.size(); ++pi) {
        const uint64_t row_lo = passes[pi].first, row_hi = passes[pi].second;
        if (row_lo >= row_hi) continue;
        tabs.pass = (int)pi;
        // 3a. this pass's candidate count and per-source offsets
        const uint64_t n_pass = pass_count[pi];
        const uint64_t* off = nullptr;
        if (passes.size() > 1) {
            uint64_t* offs = c->src_off.get<uint64_t>(ns + 1);
            exclusive_scan_u32(c, pcnt + pi * ns, ns, offs, offs + ns);
            off = offs;
            uint64_t tot = 0;
            HIPCHK(hipMemcpyAsync(&tot, offs + ns, 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if (tot != n_pass) throw HipError("per-source pass counts disagree with the planning histogram", -1);
        }
        if (n_pass == 0) continue;
        using R = typename RecOf<T>::type;
        using RC = typename RecCOf<T>::type;
        // 3b/3c. group by target and merge exactly.  Large passes: hash buckets (two
        // partition passes) merged in LDS tables, a bucket with more targets than a table
        // holds split into hash classes inside the merge kernel; small passes, or a pass whose
        // partition region or bucket classes overflow, take the full sort on the dense index
        // and the segmented reduction.
        int hb = 0;
        if (c->merge_mode >= 2) {
            hb = c->merge_hash_bits;
        } else if (c->merge_mode == 0 && n_pass >= ((size_t)1 << 20)) {
            // distinct targets expected in this pass; aim at a table load below 0.6, and drop
            // to a whole number of 8-bit radix digits when that keeps the load below 0.8
            const double est = pass_dist[pi];
            while (hb < 16 && est > 0.6 * BR_SLOTS_C * (double)(1ull << hb)) ++hb;
            int down = hb & ~7;
            if (down > 0 && down < hb && est <= 0.8 * BR_SLOTS_C * (double)(1ull << down)) hb = down;
        }
        bool hashed = hb > 0 && hb <= 16;
        // compact records (12 or 8 bytes) when the pass's dense range leaves at most 31 bits
        // below the bucket bits; automatic planning raises hb (up to 16) to get there
        const int kbits = std::max(1, bits_for((row_hi - row_lo) * H.dim));
        if (hashed && c->merge_mode == 0 && hb > 8 && kbits - hb > 31 && kbits - 31 <= 16) hb = kbits - 31;
        bool compact = hashed && hb > 8 && kbits - hb <= 31 && c->merge_mode != 4;
        CHash hp{};
        if (compact) {
            hp.kbits = kbits;
            hp.hb = hb;
            hp.rb = kbits - hb;
            hp.mask = kbits >= 64 ? ~0ull : ((1ull << kbits) - 1);
            hp.mul = (0x9E3779B97F4A7C15ull & hp.mask) | 1ull;
            uint64_t inv = hp.mul;   // Newton: inv <- inv (2 - mul inv), exact mod 2^64 after 6 steps
            for (int it = 0; it < 6; ++it) inv *= 2 - hp.mul * inv;
            hp.inv = inv & hp.mask;
        }
        // A: generation output; B: the (gapped) first-level partition, also room for the
        // key/value split of the sort path; compact passes: their buckets in cand_v[0]
        const size_t capB = part_capacity(n_pass);
        R* A = (R*)c->cand_k[0].get<uint8_t>(n_pass * sizeof(R));
        R* B = (R*)c->cand_k[1].get<uint8_t>(std::max(capB * sizeof(R), n_pass * (sizeof(uint64_t) + sizeof(T))));
        unsigned long long* d_snext = (unsigned long long*)c->d_counter + 7;   // the generator's source counter
        auto generate = [&](bool cr) {
            PhaseTimer t(c);
            HIPCHK(hipMemsetAsync(d_snext, 0, 8, c->stream));
            if (cr) {
                uint16_t* bk = c->cand_v[0].get<uint16_t>(n_pass);
                int tg = kt_begin(c, K_GENERATE, ((double)sizeof(RC) + 2.0) * n_pass + 16.0 * ns);
                k_generate_compact<T><<<grid_for(c, ns), SFCI_BLOCK, 0, c->stream>>>(
                    H, sk, sv, ns, lambda, coeff, row_lo, row_hi, off, c->d_sdec, CompactOut{A, bk, hp}, n_pass,
                    dterm, c->d_flags, tabs, d_snext);
                kt_end(c, tg);
            } else {
                int tg = kt_begin(c, K_GENERATE, (double)sizeof(R) * n_pass + 16.0 * ns);
                k_generate<T><<<grid_for(c, ns), SFCI_BLOCK, 0, c->stream>>>(H, sk, sv, ns, lambda, coeff, row_lo,
                                                                            row_hi, off, c->d_sdec, A, n_pass,
                                                                            dterm, c->d_flags, tabs, d_snext);
                kt_end(c, tg);
            }
            HIPCHK(hipGetLastError());
            t.lap(&st.ms_generate);
        };
        generate(compact);
        st.n_candidates += n_pass;
        R* rd = A;        // the records in their current order
        R* other = B;     // the free buffer
        uint64_t* bstart = nullptr;
        uint64_t* bend = nullptr;
        if (hashed) {
            PhaseTimer t(c);
            const size_t nb = (size_t)1 << hb;
            bstart = c->bucket_bounds.get<uint64_t>(2 * nb);
            bend = bstart + nb;
            if (compact) {
                if (!hash_partition_c<RC>(c, (RC*)A, c->cand_v[0].get<uint16_t>(n_pass), (RC*)B, n_pass, hb, bstart,
                                          bend)) {
                    // a region overflowed: the pass is generated again with full keys and sorted
                    hashed = compact = false;
                    st.n_fallback += 1;
                    generate(false);
                }
            } else {
                R* g = hash_partition<R>(c, A, B, n_pass, hb);
                if (!g) {
                    hashed = false;
                    st.n_fallback += 1;
                } else {
                    rd = g;
                    other = (g == A) ? B : A;
                    int tb = kt_begin(c, K_BUCKET_BOUNDS, 16.0 * nb);
                    k_bucket_bounds<R><<<(unsigned)((nb + SFCI_BLOCK - 1) / SFCI_BLOCK), SFCI_BLOCK, 0, c->stream>>>(
                        rd, n_pass, hb, nb, bstart, bend);
                    kt_end(c, tb);
                }
            }
            t.lap(&st.ms_sort);
        }
        // sort path: split the records into keys and values in the free buffer, sort them
        // with the record buffer as scratch
        uint64_t* kd = nullptr;
        T* vd = nullptr;
        auto sort_path = [&]() {
            PhaseTimer ts(c);
            kd = (uint64_t*)other;
            vd = (T*)(kd + n_pass);
            uint64_t* kt = (uint64_t*)rd;
            T* vt = (T*)(kt + n_pass);
            k_rec_split<R, T><<<grid_for(c, (n_pass + SFCI_BLOCK - 1) / SFCI_BLOCK, 16), SFCI_BLOCK, 0, c->stream>>>(
                rd, n_pass, kd, vd);
            int bits = bits_for(2 * (row_hi - row_lo) * H.dim);
            radix_sort_kv<T>(c, &kd, &vd, kt, vt, n_pass, bits);
            ts.lap(&st.ms_sort);
        };
        if (!hashed) sort_path();
        unsigned long long* d_split = (unsigned long long*)c->d_counter + 3;
        unsigned long long* d_bnext = (unsigned long long*)c->d_counter + 6;   // the merge's bucket counter
        for (int attempt = 0;; ++attempt) {
            PhaseTimer t(c);
            uint64_t before = out_count_host;
            acc_clear(c, 1);
            HIPCHK(hipMemsetAsync(d_split, 0, 8, c->stream));
            HIPCHK(hipMemsetAsync(d_bnext, 0, 8, c->stream));
            if (hashed) {
                const uint64_t nb = (uint64_t)1 << hb;
                // (mode 3, a test knob: every table counts as full at BR_FILL / 16 slots)
                const uint32_t fill = c->merge_mode == 3 ? BR_FILL / 16 : compact ? BR_FILL_C : BR_FILL;
                if (compact) {
                    int tr = kt_begin(c, K_BUCKET_REDUCE, (double)sizeof(RC) * n_pass);
                    k_bucket_reduce<T, true><<<256, BR_THREADS, 0, c->stream>>>(
                        H, A, hp, bstart, bend, nb, hb, row_lo * H.dim, sk, sv, dterm, ns, cut_eff, stk, stv,
                        d_out_count, stage_cap, d_uniq, c->d_acc + 2, c->d_flags, d_split, fill, d_bnext);
                    kt_end(c, tr);
                } else {
                    int tr = kt_begin(c, K_BUCKET_REDUCE, (double)sizeof(R) * n_pass);
                    k_bucket_reduce<T, false><<<256, BR_THREADS, 0, c->stream>>>(
                        H, rd, hp, bstart, bend, nb, hb, row_lo * H.dim, sk, sv, dterm, ns, cut_eff, stk, stv,
                        d_out_count, stage_cap, d_uniq, c->d_acc + 2, c->d_flags, d_split, fill, d_bnext);
                    kt_end(c, tr);
                }
            } else {
                int grid = (int)std::min<size_t>((n_pass + RED_TILE - 1) / RED_TILE, 256 * 8);
                int tr = kt_begin(c, K_REDUCE, (8.0 + sizeof(T)) * n_pass);
                k_reduce<T><<<grid, SFCI_BLOCK, 0, c->stream>>>(H, kd, vd, n_pass, row_lo * H.dim, sk, sv, dterm, ns,
                                                                cut_eff, stk, stv, d_out_count, stage_cap, d_uniq,
                                                                c->d_acc + 2, c->d_flags);
                kt_end(c, tr);
            }
            HIPCHK(hipGetLastError());
            uint64_t cnts[3];
            HIPCHK(hipMemcpyAsync(cnts, d_out_count, 16, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(cnts + 2, d_split, 8, hipMemcpyDeviceToHost, c->stream));
            t.lap(&st.ms_reduce);
            HIPCHK(hipStreamSynchronize(c->stream));
            if (hashed) {
                const uint32_t fl_now = flags_read_clear(c);
                if (fl_now & FLAG_BUCKET_FULL) {
                    // a class still overflowed at 256 classes per bucket: roll back this pass's
                    // output and merge it through the full sort instead
                    st.n_fallback += 1;
                    uint64_t restore[2] = {before, uniq_host};
                    HIPCHK(hipMemcpyAsync(d_out_count, restore, 16, hipMemcpyHostToDevice, c->stream));
                    if (fl_now & ~(uint32_t)(FLAG_BUCKET_FULL | FLAG_OUT_OVERFLOW)) atomic_flags_restore(c, fl_now);
                    if (compact) {
                        // compact records carry no full key: generate the pass again with them
                        compact = false;
                        rd = A;
                        other = B;
                        generate(false);
                    }
                    hashed = false;
                    sort_path();
                    continue;
                }
                if (fl_now & ~(uint32_t)FLAG_OUT_OVERFLOW) atomic_flags_restore(c, fl_now);
            }
            if (cnts[0] <= stage_cap) {
                out_count_host = cnts[0];
                rr_lo.push_back(row_lo);
                rr_ratio.push_back(std::max(1e-6, (double)(cnts[1] - uniq_host) / (double)n_pass));
                uniq_host = cnts[1];
                st.n_split += cnts[2];
                if (hashed) buckets_merged += (uint64_t)1 << hb;
                diag_total = fx_add(diag_total, acc_read_raw(c, 1));
                break;
            }
            if (attempt > 2) throw HipError("merged-output staging overflow after regrow", -3);
            // grow the staging list, keep earlier passes' output, redo this pass's merge
            size_t ncap = cnts[0] + (cnts[0] >> 2);
            DevBuf nk, nv;
            uint64_t* nkp = nk.get<uint64_t>(ncap);
            double* nvp = nv.get<double>(ncap);
            if (before) {
                HIPCHK(hipMemcpyAsync(nkp, stk, before * 8, hipMemcpyDeviceToDevice, c->stream));
                HIPCHK(hipMemcpyAsync(nvp, stv, before * 8, hipMemcpyDeviceToDevice, c->stream));
            }
            HIPCHK(hipStreamSynchronize(c->stream));
            c->stage_k.release();
            c->stage_v.release();
            c->stage_k = nk;
            c->stage_v = nv;
            stk = nkp;
            stv = nvp;
            stage_cap = ncap;
            cnts[0] = before;
            cnts[1] = uniq_host;
            HIPCHK(hipMemcpyAsync(d_out_count, cnts, 16, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            uint32_t fl = flags_read_clear(c);
            if (fl & FLAG_RANGE) throw HipError("a term left the fixed-point range (|x| >= 2^42)", -5);
        }
        st.hash_bits = hashed ? hb : 0;
    }
    uint64_t uniq = uniq_host;
    *diag = fx_to(diag_total);
    uint32_t fl = flags_read_clear(c);
    if (fl & FLAG_RANGE) throw HipError("a term left the fixed-point range (|x| >= 2^42)", -5);
    if (fl & FLAG_GEN_OVERFLOW) throw HipError("generation exceeded the planning count", -1);
    if (fl & FLAG_OUT_OVERFLOW) throw HipError("merged-output staging overflow", -3);
    if (fl & FLAG_BAD_INPUT) throw HipError("input vector keys are not sorted and unique", -1);
    st.n_unique = uniq;
    if (st.n_candidates) c->unique_ratio = std::max(1e-6, (double)uniq / (double)st.n_candidates);
    if (rr_lo.size() > 1 && rr_lo[0] == 0) {
        c->rr_lo.swap(rr_lo);
        c->rr_ratio.swap(rr_ratio);
    } else {
        c->rr_lo.clear();
        c->rr_ratio.clear();
    }
    if (c->merge_mode == 0 && buckets_merged) {
        // the distinct-target count of a pass varies with its row range: aim lower while more
        // than 1% of the buckets split, back up while none does
        const double split_frac = (double)st.n_split / (double)buckets_merged;
        if (split_frac > 0.01) c->pass_load = std::max(0.5, c->pass_load * 0.85);
        else if (st.n_split == 0) c->pass_load = std::min(0.9, c->pass_load * 1.05);
    }
    // 4. staged dense indices -> determinant keys; top out_cap by |y|, then key order
    size_t m = out_count_host;
    if (m > 0) {
        k_dense_to_det<<<grid_for(c, (m + SFCI_BLOCK - 1) / SFCI_BLOCK, 16), SFCI_BLOCK, 0, c->stream>>>(H, stk, m);
        HIPCHK(hipGetLastError());
    }
    size_t kept = 0;
    if (all) {
        PhaseTimer t(c);
        uint64_t* tk = c->tmp_k[1].get<uint64_t>(std::max<size_t>(m, 1));
        double* tv = c->tmp_v[1].get<double>(std::max<size_t>(m, 1));
        uint64_t* kk = stk;
        double* vv = stv;
        radix_sort_pairs(c, &kk, &vv, tk, tv, m, c->n_so);
        all->k = kk;
        all->v = vv;
        all->n = m;
        kept = m;
        t.lap(&st.ms_select);
    } else if (out_cap > 0 && m > 0) {
        PhaseTimer t(c);
        uint64_t* sel_k = c->tmp_k[0].get<uint64_t>(m);
        double* sel_v = c->tmp_v[0].get<double>(m);
        kept = select_top(c, stk, stv, m, out_cap, sel_k, sel_v);
        uint64_t* tk = c->tmp_k[1].get<uint64_t>(kept);
        double* tv = c->tmp_v[1].get<double>(kept);
        uint64_t* kk = sel_k;
        double* vv = sel_v;
        radix_sort_pairs(c, &kk, &vv, tk, tv, kept, c->n_so);
        dev_copy(c, ok, ov, kk, vv, kept);
        t.lap(&st.ms_select);
    }
    st.n_out = kept;
    total.lap(&st.ms_total);
    kt_resolve(c);
    return kept;
}

size_t dev_hv(Ctx* c, const uint64_t* ik, const double* iv, size_t n_in, size_t nd, double lambda, double coeff,
              size_t out_cap, uint64_t* ok, double* ov, double* diag) {
    if (c->precision == 1) return dev_hv_t<float>(c, ik, iv, n_in, nd, lambda, coeff, out_cap, ok, ov, diag, nullptr);
    return dev_hv_t<double>(c, ik, iv, n_in, nd, lambda, coeff, out_cap, ok, ov, diag, nullptr);
}

__global__ void k_fill(double* __restrict__ v, size_t n, double x) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v[i] = x;
}

void dev_connection_records(Ctx* c, const uint64_t* sk, size_t ns, const Rec64** rec, const double** dterm) {
    const HamDev& H = c->ham;
    const size_t n = ns * H.n_conn;
    double* ones = c->src_v.get<double>(ns);
    Rec64* r0 = (Rec64*)c->cand_k[0].get<uint8_t>(n * sizeof(Rec64));
    double* dt = c->dterm.get<double>(ns);
    k_fill<<<grid_for(c, ns), SFCI_BLOCK, 0, c->stream>>>(ones, ns, 1.0);
    k_generate<double><<<grid_for(c, ns), SFCI_BLOCK, 0, c->stream>>>(H, sk, ones, ns, 0.0, 1.0, 0, H.dim, nullptr,
                                                                      c->d_sdec, r0, n, dt, c->d_flags, SsTabs{},
                                                                      nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint32_t fl = flags_read_clear(c);
    if (fl & FLAG_RANGE) throw HipError("a term left the fixed-point range (|x| >= 2^42)", -5);
    if (fl & FLAG_GEN_OVERFLOW) throw HipError("generation exceeded its buffer", -1);
    *rec = r0;
    *dterm = dt;
}


double dev_energy(Ctx* c, const uint64_t* k, const double* v, size_t n) {
    const HamDev& H = c->ham;
    if (n == 0) return 0.0;
    if (n >= 0xffffffffull) throw HipError("energy: vector longer than 2^32 - 1", -4);
    uint64_t* dk = c->tmp_k[0].get<uint64_t>(n);
    double* dv = c->tmp_v[0].get<double>(n);
    uint64_t* dk2 = c->tmp_k[1].get<uint64_t>(n);
    double* dv2 = c->tmp_v[1].get<double>(n);
    k_dense_split<<<grid_for(c, n, 16), SFCI_BLOCK, 0, c->stream>>>(H, k, v, n, dk, dv);
    uint64_t* sk = dk;
    double* sv = dv;
    radix_sort_pairs(c, &sk, &sv, dk2, dv2, n, bits_for(H.dim * H.dim));
    uint32_t* row_off = c->sel_b.get<uint32_t>(H.dim + 1);
    uint32_t* r1s = c->sel_a.get<uint32_t>(n);
    k_row_index<<<grid_for(c, std::max<size_t>(n, H.dim + 1), 16), SFCI_BLOCK, 0, c->stream>>>(sk, n, H.dim, row_off,
                                                                                                r1s);
    acc_clear(c, 4);
    unsigned long long* d_snext = (unsigned long long*)c->d_counter + 7;   // the pull's source counter
    HIPCHK(hipMemsetAsync(d_snext, 0, 8, c->stream));
    int te = kt_begin(c, K_ENERGY, 16.0 * n);
    k_energy<<<grid_for(c, n), SFCI_BLOCK, 0, c->stream>>>(H, k, v, n, row_off, r1s, sv, c->d_sdec, c->d_acc + 8,
                                                           c->d_flags, d_snext);
    kt_end(c, te);
    HIPCHK(hipGetLastError());
    double e = acc_read(c, 4);
    if (flags_read_clear(c) & FLAG_RANGE) throw HipError("a term left the fixed-point range (|x| >= 2^42)", -5);
    kt_resolve(c);
    return e;
}

void dev_hv_all(Ctx* c, const uint64_t* ik, const double* iv, size_t n_in, double lambda, double coeff, double* diag,
                HvAll* all) {
    *all = HvAll{};
    if (c->precision == 1) dev_hv_t<float>(c, ik, iv, n_in, n_in, lambda, coeff, 0, nullptr, nullptr, diag, all);
    else dev_hv_t<double>(c, ik, iv, n_in, n_in, lambda, coeff, 0, nullptr, nullptr, diag, all);
}


// ---------------------------------------------------------------------------------
// Screened H.v.  One pass over all target rows:
//   generate  every connection evaluated as in the mixed-precision path, only those with
//             |H_ij coeff c_j| > eps (and every diagonal tag) written (k_generate<float, true>)
//   score     the network's prediction of log|y_t| (mode 2) or |value| (mode 1)
//   select    the candidates whose score reaches the keep-th largest (ties kept), by two
//             16-bit histogram rounds over an order-preserving integer form of the score
//   merge     key sort + the exact segmented reduction, then the cut, cap and key order
// The kept set depends only on the candidates' contents, and the merge is exact, so the
// result does not depend on the order in which the generation appended them.

// order-preserving map of a float to uint32 (larger float -> larger integer)
__device__ __forceinline__ uint32_t score_bits(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ void k_score_mag(const RecS* __restrict__ r, size_t n, float* __restrict__ score) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const RecS x = r[i];
        score[i] = (x.key & 1ull) ? INFINITY : fabsf(x.val);
    }
}

// Histogram of bits [shift, shift + width) of score_bits over the scores whose bits above
// shift + width equal `prefix` (all of them when shift + width == 32), privatized in LDS.
constexpr int SEL_BITS = 11;
__global__ __launch_bounds__(SFCI_BLOCK) void k_score_hist(const float* __restrict__ score, size_t n, int shift,
                                                           int width, uint32_t prefix, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[1 << SEL_BITS];
    const uint32_t nb = 1u << width;
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const int top = shift + width;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t u = score_bits(score[i]);
        if (top >= 32 || (u >> top) == prefix) atomicAdd(&h[(u >> shift) & (nb - 1)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
        if (h[i]) atomicAdd(hist + i, h[i]);
}

// kept records (score >= threshold) as split key / value arrays, appended with one atomic
// per wavefront
__global__ __launch_bounds__(SFCI_BLOCK) void k_keep_compact(const RecS* __restrict__ r, const float* __restrict__ score,
                                                             size_t n, uint32_t thr, uint64_t* __restrict__ kout,
                                                             float* __restrict__ vout,
                                                             unsigned long long* __restrict__ count) {
    const int lane = threadIdx.x & (SFCI_WAVE - 1);
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i0 = (size_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
        const size_t i = i0 + threadIdx.x;
        const bool keep = i < n && score_bits(score[i]) >= thr;
        const uint64_t bal = __ballot(keep);
        unsigned long long base = 0;
        if (lane == 0 && bal) base = atomicAdd(count, (unsigned long long)popc64(bal));
        base = __shfl(base, 0, SFCI_WAVE);
        if (keep) {
            const uint64_t at = base + popc64(bal & lt);
            kout[at] = r[i].key;
            vout[at] = r[i].val;
        }
    }
}

// all records, split (nothing screened out)
__global__ void k_recs_split(const RecS* __restrict__ r, size_t n, uint64_t* __restrict__ kout,
                             float* __restrict__ vout) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        kout[i] = r[i].key;
        vout[i] = r[i].val;
    }
}

namespace {

// Screened generation into c->cand_k[0]; returns the record count.  A list that would not
// fit the workspace raises eps fourfold and runs again.
size_t screened_generate(Ctx* c, const uint64_t* sk, const double* sv, size_t ns, double lambda, double coeff,
                         double* dterm, double eps, RecS** recs_out, double* eps_used) {
    const HamDev& H = c->ham;
    const uint64_t total = (uint64_t)ns * H.n_conn;
    const size_t cap = (size_t)std::max<uint64_t>(1, std::min<uint64_t>(total, c->ws_limit / (2 * sizeof(RecS))));
    RecS* recs = (RecS*)c->cand_k[0].get<uint8_t>(cap * sizeof(RecS));
    unsigned long long* cnt = (unsigned long long*)c->d_counter + 2;
    for (int attempt = 0;; ++attempt) {
        HIPCHK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), c->stream));
        unsigned long long* d_snext = (unsigned long long*)c->d_counter + 7;   // the generator's source counter
        HIPCHK(hipMemsetAsync(d_snext, 0, 8, c->stream));
        int tg = kt_begin(c, K_GENERATE, 16.0 * ns);
        k_generate_screened<<<grid_for(c, ns), SFCI_BLOCK, 0, c->stream>>>(
            H, sk, sv, ns, lambda, coeff, c->d_sdec, dterm, c->d_flags, ScreenOut{recs, cnt, (uint64_t)cap, (float)eps},
            d_snext);
        kt_end(c, tg);
        HIPCHK(hipGetLastError());
        unsigned long long m = 0;
        HIPCHK(hipMemcpyAsync(&m, cnt, sizeof(m), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        const uint32_t fl = flags_read_clear(c);
        if (fl & FLAG_RANGE) throw HipError("a term left the fixed-point range (|x| >= 2^42)", -5);
        if (fl & FLAG_BAD_INPUT) throw HipError("input vector keys are not sorted and unique", -1);
        if (!(fl & FLAG_GEN_OVERFLOW) && m <= cap) {
            *recs_out = recs;
            *eps_used = eps;
            return (size_t)m;
        }
        if (attempt >= 8) throw HipError("screened candidates exceed the workspace at every prefilter tried", -4);
        eps = eps > 0.0 ? eps * 4.0 : 1e-12;
    }
}

// the largest threshold u with #(score_bits >= u) >= keep, by three histogram rounds
// (bits 31..21, 20..10, 9..0)
uint32_t score_threshold(Ctx* c, const float* score, size_t m, size_t keep) {
    uint32_t* hist = c->scratch_hist.get<uint32_t>(1 << SEL_BITS);
    std::vector<uint32_t> h(1 << SEL_BITS);
    const int g = grid_for(c, (m + SFCI_BLOCK - 1) / SFCI_BLOCK, 8);
    uint64_t above = 0;
    uint32_t prefix = 0;
    const int shifts[3] = {21, 10, 0}, widths[3] = {11, 11, 10};
    for (int round = 0; round < 3; ++round) {
        const int nb = 1 << widths[round];
        HIPCHK(hipMemsetAsync(hist, 0, nb * sizeof(uint32_t), c->stream));
        k_score_hist<<<g, SFCI_BLOCK, 0, c->stream>>>(score, m, shifts[round], widths[round], prefix, hist);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(h.data(), hist, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        int b = nb - 1;
        for (; b > 0; --b) {
            if (above + h[b] >= keep) break;
            above += h[b];
        }
        prefix = prefix << widths[round] | (uint32_t)b;
    }
    return prefix;
}

}  // namespace

size_t dev_hv_screened(Ctx* c, const uint64_t* ik, const double* iv, size_t n_in, size_t nd, double lambda,
                       double coeff, size_t out_cap, uint64_t* ok, double* ov, double* diag) {
    const HamDev& H = c->ham;
    auto& st = c->st;
    kt_resolve(c);   // laps left pending by an earlier call that threw
    st = {};
    PhaseTimer total(c);
    *diag = 0.0;
    if (n_in == 0) return 0;
    if (c->screen_mode >= 2 && !c->screen_net) throw HipError("network screening without a network", -1);
    const uint64_t* sk = ik;
    const double* sv = iv;
    size_t ns = n_in;
    if (n_in > nd) {
        uint64_t* tk = c->src_k.get<uint64_t>(nd);
        double* tv = c->src_v.get<double>(nd);
        ns = select_top(c, ik, iv, n_in, nd, tk, tv);
        sk = tk;
        sv = tv;
    }
    st.n_src = ns;
    st.n_passes = 1;
    double* dterm = c->dterm.get<double>(ns);
    // 1. screened generation
    RecS* recs = nullptr;
    double eps_used = 0.0;
    size_t m = 0;
    {
        PhaseTimer t(c);
        m = screened_generate(c, sk, sv, ns, lambda, coeff, dterm, c->screen_eps, &recs, &eps_used);
        t.lap(&st.ms_generate);
    }
    st.n_candidates = m;
    st.screen_eps = eps_used;
    // 2. score and select
    const size_t keep = std::max<size_t>((size_t)(c->screen_keep * (double)nd), ns + 1);
    // keys + values of the kept records, and their sort scratch, in the second buffer
    const size_t m_alloc = std::max<size_t>(m, 1);
    uint8_t* buf = c->cand_k[1].get<uint8_t>(m_alloc * (2 * (sizeof(uint64_t) + sizeof(float))) + m_alloc * sizeof(float));
    uint64_t* kd = (uint64_t*)buf;
    uint64_t* kt = kd + m_alloc;
    float* vd = (float*)(kt + m_alloc);
    float* vt = vd + m_alloc;
    float* score = vt + m_alloc;
    size_t kept = m;
    if (m > keep) {
        PhaseTimer t(c);
        if (c->screen_mode >= 2) {
            mlp_score_records(c, c->screen_net, recs, m, score, c->screen_mode == 3);
        } else {
            k_score_mag<<<grid_for(c, (m + SFCI_BLOCK - 1) / SFCI_BLOCK, 16), SFCI_BLOCK, 0, c->stream>>>(recs, m, score);
        }
        HIPCHK(hipGetLastError());
        t.lap(&st.ms_score);
        PhaseTimer t2(c);
        int ts = kt_begin(c, K_SCREEN_SELECT, 20.0 * m + 12.0 * keep);
        const uint32_t thr = score_threshold(c, score, m, keep);
        unsigned long long* cnt = (unsigned long long*)c->d_counter + 3;
        HIPCHK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), c->stream));
        k_keep_compact<<<grid_for(c, (m + SFCI_BLOCK - 1) / SFCI_BLOCK, 16), SFCI_BLOCK, 0, c->stream>>>(
            recs, score, m, thr, kd, vd, cnt);
        HIPCHK(hipGetLastError());
        unsigned long long kc = 0;
        HIPCHK(hipMemcpyAsync(&kc, cnt, sizeof(kc), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        kt_end(c, ts);
        kept = (size_t)kc;
        t2.lap(&st.ms_select);
    } else if (m > 0) {
        k_recs_split<<<grid_for(c, (m + SFCI_BLOCK - 1) / SFCI_BLOCK, 16), SFCI_BLOCK, 0, c->stream>>>(recs, m, kd, vd);
        HIPCHK(hipGetLastError());
    }
    st.n_kept = kept;
    // 3. exact merge of the kept candidates
    size_t stage_cap = std::max<size_t>(std::max<size_t>(out_cap + (out_cap >> 2), kept + 1), (size_t)1 << 16);
    if (c->stage_k.bytes / sizeof(uint64_t) > stage_cap) stage_cap = c->stage_k.bytes / sizeof(uint64_t);
    uint64_t* stk = c->stage_k.get<uint64_t>(stage_cap);
    double* stv = c->stage_v.get<double>(stage_cap);
    const double cut_eff = out_cap ? c->cut : INFINITY;
    unsigned long long* d_out_count = (unsigned long long*)c->d_counter;
    unsigned long long* d_uniq = d_out_count + 1;
    HIPCHK(hipMemsetAsync(c->d_counter, 0, 2 * sizeof(uint64_t), c->stream));
    uint64_t cnts[2] = {0, 0};
    if (kept > 0) {
        PhaseTimer ts(c);
        radix_sort_kv<float>(c, &kd, &vd, kt, vt, kept, bits_for(2 * H.dim * H.dim));
        ts.lap(&st.ms_sort);
        PhaseTimer t(c);
        acc_clear(c, 1);
        int grid = (int)std::min<size_t>((kept + RED_TILE - 1) / RED_TILE, 256 * 8);
        int tr = kt_begin(c, K_REDUCE, 12.0 * kept);
        k_reduce<float><<<grid, SFCI_BLOCK, 0, c->stream>>>(H, kd, vd, kept, 0, sk, sv, dterm, ns, cut_eff, stk, stv,
                                                            d_out_count, stage_cap, d_uniq, c->d_acc + 2, c->d_flags);
        kt_end(c, tr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(cnts, d_out_count, 16, hipMemcpyDeviceToHost, c->stream));
        *diag = acc_read(c, 1);
        t.lap(&st.ms_reduce);
    }
    const uint32_t fl = flags_read_clear(c);
    if (fl & FLAG_RANGE) throw HipError("a term left the fixed-point range (|x| >= 2^42)", -5);
    if (fl & FLAG_OUT_OVERFLOW) throw HipError("merged-output staging overflow", -3);
    if (fl & FLAG_BAD_INPUT) throw HipError("input vector keys are not sorted and unique", -1);
    st.n_unique = cnts[1];
    // 4. output: determinant keys, top out_cap by |y|, key order
    const size_t mo = (size_t)cnts[0];
    if (mo > 0) {
        k_dense_to_det<<<grid_for(c, (mo + SFCI_BLOCK - 1) / SFCI_BLOCK, 16), SFCI_BLOCK, 0, c->stream>>>(H, stk, mo);
        HIPCHK(hipGetLastError());
    }
    size_t n_out = 0;
    if (out_cap > 0 && mo > 0) {
        PhaseTimer t(c);
        uint64_t* sel_k = c->tmp_k[0].get<uint64_t>(mo);
        double* sel_v = c->tmp_v[0].get<double>(mo);
        n_out = select_top(c, stk, stv, mo, out_cap, sel_k, sel_v);
        uint64_t* tk = c->tmp_k[1].get<uint64_t>(n_out);
        double* tv = c->tmp_v[1].get<double>(n_out);
        uint64_t* kk = sel_k;
        double* vv = sel_v;
        radix_sort_pairs(c, &kk, &vv, tk, tv, n_out, c->n_so);
        dev_copy(c, ok, ov, kk, vv, n_out);
        t.lap(&st.ms_select);
    }
    st.n_out = n_out;
    total.lap(&st.ms_total);
    kt_resolve(c);
    return n_out;
}

// ---- training samples ---------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64s(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t rec_hash(const RecS& r, uint64_t seed) {
    return mix64s(seed ^ mix64s(r.key ^ ((uint64_t)__float_as_uint(r.val) << 32 | r.cls)));
}

// (hash, record index) of the off-diagonal records whose hash is below thr
__global__ __launch_bounds__(SFCI_BLOCK) void k_sample_pick(const RecS* __restrict__ r, size_t n, uint64_t seed,
                                                            uint64_t thr, uint64_t* __restrict__ hk,
                                                            double* __restrict__ idx,
                                                            unsigned long long* __restrict__ count) {
    const int lane = threadIdx.x & (SFCI_WAVE - 1);
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i0 = (size_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
        const size_t i = i0 + threadIdx.x;
        uint64_t h = 0;
        bool take = false;
        if (i < n) {
            const RecS x = r[i];
            h = rec_hash(x, seed);
            take = !(x.key & 1ull) && h < thr;
        }
        const uint64_t bal = __ballot(take);
        unsigned long long base = 0;
        if (lane == 0 && bal) base = atomicAdd(count, (unsigned long long)popc64(bal));
        base = __shfl(base, 0, SFCI_WAVE);
        if (take) {
            const uint64_t at = base + popc64(bal & lt);
            hk[at] = h;
            const uint64_t ii = i;
            double d;
            __builtin_memcpy(&d, &ii, 8);
            idx[at] = d;
        }
    }
}

// samples in hash order, labelled by a binary search of the target in the full product
__global__ void k_sample_fill(HamDev H, const RecS* __restrict__ r, const double* __restrict__ idx, size_t n,
                              const uint64_t* __restrict__ yk, const double* __restrict__ yv, size_t ny,
                              sfci_screen_sample_t* __restrict__ out, uint32_t* flags) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint64_t ii;
        const double d = idx[i];
        __builtin_memcpy(&ii, &d, 8);
        const RecS x = r[ii];
        const uint64_t det = x.det;
        size_t lo = 0, hi = ny;
        while (lo < hi) {
            const size_t mid = (lo + hi) >> 1;
            if (yk[mid] < det) lo = mid + 1;
            else hi = mid;
        }
        double y = 0.0;
        if (lo < ny && yk[lo] == det) y = yv[lo];
        else atomicOr(flags, FLAG_BAD_INPUT);
        sfci_screen_sample_t s;
        s.det = det;
        s.logv = plog2(x.val);
        s.label = (float)log(y != 0.0 ? fabs(y) : 1e-30);
        s.cls = x.cls;
        s.pad = 0;
        out[i] = s;
    }
}

size_t dev_screen_samples(Ctx* c, const uint64_t* vk, const double* vv, size_t n, double eps, size_t max_samples,
                          uint64_t seed, sfci_screen_sample_t* out) {
    if (n == 0 || max_samples == 0) return 0;
    // full uncut product: every distinct target, key-sorted (library buffers)
    HvAll all;
    double d = 0.0;
    dev_hv_all(c, vk, vv, n, 0.0, 1.0, &d, &all);
    // screened candidates of the same product
    double* dterm = c->dterm.get<double>(n);
    RecS* recs = nullptr;
    double eps_used = 0.0;
    const size_t m = screened_generate(c, vk, vv, n, 0.0, 1.0, dterm, eps, &recs, &eps_used);
    if (eps_used != eps) throw HipError("sample candidates exceed the workspace at this prefilter", -4);
    if (m == 0) return 0;
    const double frac = std::min(1.0, 1.25 * (double)max_samples / (double)m);
    const uint64_t thr = frac >= 1.0 ? ~0ull : (uint64_t)(frac * 18446744073709551615.0);
    uint64_t* hk = c->sel_a.get<uint64_t>(2 * m);
    double* hx = (double*)(hk + m);
    uint64_t* hk2 = c->sel_b.get<uint64_t>(2 * m);
    double* hx2 = (double*)(hk2 + m);
    unsigned long long* cnt = (unsigned long long*)c->d_counter + 3;
    HIPCHK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), c->stream));
    k_sample_pick<<<grid_for(c, (m + SFCI_BLOCK - 1) / SFCI_BLOCK, 16), SFCI_BLOCK, 0, c->stream>>>(recs, m, seed, thr,
                                                                                                   hk, hx, cnt);
    HIPCHK(hipGetLastError());
    unsigned long long np = 0;
    HIPCHK(hipMemcpyAsync(&np, cnt, sizeof(np), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (np == 0) return 0;
    radix_sort_pairs(c, &hk, &hx, hk2, hx2, (size_t)np, 64);
    const size_t take = std::min<size_t>((size_t)np, max_samples);
    DevBuf ob;
    sfci_screen_sample_t* dout = ob.get<sfci_screen_sample_t>(take);
    k_sample_fill<<<grid_for(c, (take + SFCI_BLOCK - 1) / SFCI_BLOCK, 16), SFCI_BLOCK, 0, c->stream>>>(
        c->ham, recs, hx, take, all.k, all.v, all.n, dout, c->d_flags);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dout, take * sizeof(sfci_screen_sample_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    ob.release();
    const uint32_t fl = flags_read_clear(c);
    if (fl & FLAG_BAD_INPUT) throw HipError("a screened candidate's target is missing from the full product", -1);
    return take;
}

}  // namespace sfci
