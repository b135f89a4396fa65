// CPU restatement of the reference SFCI hot path (wlj89/SFCI) -- TEST INFRASTRUCTURE.
//
// This library is the oracle the GPU library is checked against and the CPU baseline
// that bench.py times.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg load it; the product never does.
//
// What it restates (reference file:line):
//   ReadElements / load_integrals        src/functions.cpp:297-418
//   indexCal / nck / nck_list            src/bitstr.cpp:54-131, src/main.cpp:183-223
//   signCal_bit                          src/bitstr.cpp:133-160
//   OffDiagGen_bit(_e)                   src/bitstr.cpp:162-431
//   DiagCal_bit                          src/bitstr.cpp:433-476
//   merge_bucket                         src/functions.cpp:429-464
//   SpMVMtply_bucket                     src/functions.cpp:511-804
//   dynVec dot/addTwo/shrinkFrom/norm    src/all.h:249-480
//   calc_energy                          src/functions.cpp:1272-1313
//   sparse_krylov_solve_eigenval         src/sparse_solver.cpp:246-548
//
// Two H.v modes:
//   mode 0 ("exact")   -- unique emission (one candidate per connected determinant
//                         with its full Slater-Condon element), exact fixed-point
//                         merge.  The GPU library must match this bit for bit.
//   mode 1 ("literal") -- the reference's own emission (one candidate per integral
//                         list entry, OffDiagGen_bit order), stable key sort and
//                         sequential double merge.  Used to check mode 0 against the
//                         reference's arithmetic within 1e-12 relative.
//   mode 2 ("mixed")   -- mode 0 with each off-diagonal candidate formed in single
//                         precision, (float)H_ij * (float)(coeff c_j), then widened;
//                         the diagonal term stays (E_jj - lambda) * coeff c_j in double.
//                         The GPU library's mixed-precision H.v must match it bit for bit.
// Tie rule for every |c| truncation: larger |c| first, equal |c| -> smaller key first.
// The reference's introsort leaves ties unordered (src/introsort.cpp:571-580).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>
#include <omp.h>

#include "fixacc.h"

using namespace oracle;
typedef uint64_t det_t;

namespace {

struct Ent4 { uint8_t m, n, p, q; double c; };
struct Ent2 { uint8_t m, n; double c; };

struct Ham {
    int n_so = 0;
    double constant = 0.0;
    // the reference's lists (src/all.h:105-125)
    std::vector<std::vector<Ent4>> OO;   // keyed n + n_so*q
    std::vector<std::vector<Ent2>> O;    // keyed n (annihilated)
    std::vector<Ent4> OO_diag;
    std::vector<Ent2> O_diag;
    // dense spin-orbital tensors holding the same entries (duplicates summed)
    std::vector<double> G;   // ((m*N+n)*N+p)*N+q
    std::vector<double> h;   // m*N+n
    // rank tables (src/main.cpp:183-223)
    int n_alpha = 0;
    std::vector<std::vector<uint64_t>> nck_list;   // [i][j] = C(j, i+1)
    uint64_t block_dim = 0;
    unsigned bucket_bits = 0;
    uint64_t bucket_num = 0;
    double g(int m, int n, int p, int q) const {
        size_t N = n_so;
        return G[((m * N + n) * N + p) * N + q];
    }
};

static uint64_t binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    unsigned __int128 r = 1;
    for (int i = 0; i < k; ++i) { r = r * (n - i) / (i + 1); }
    return (uint64_t)r;
}

// signCal_bit (src/bitstr.cpp:133-160): (-1)^(#occupied strictly between p and q), p < q.
// D is det_t (64-bit keys) or u128 (128-bit keys, more than 64 spin orbitals).
typedef unsigned __int128 u128;
static inline int popc(uint64_t x) { return __builtin_popcountll(x); }
static inline int popc(u128 x) { return __builtin_popcountll((uint64_t)x) + __builtin_popcountll((uint64_t)(x >> 64)); }
template <class D>
static inline double sgn_between(D d, int p, int q) {
    if (p > q) std::swap(p, q);
    D mask = (q - p > 1) ? ((((D)1 << (q - p - 1)) - 1) << (p + 1)) : 0;
    return (popc(d & mask) & 1) ? -1.0 : 1.0;
}

// indexCal (src/bitstr.cpp:74-131): colex rank of the even-bit string.
static inline uint64_t alpha_rank(const Ham& H, det_t d) {
    det_t a = d & 0x5555555555555555ULL;
    uint64_t r = 0;
    int pos = 0;
    while (a) {
        int orb = __builtin_ctzll(a) >> 1;
        r += H.nck_list[pos][orb];
        a &= a - 1;
        ++pos;
    }
    return r;
}

template <class D>
static int occ_list(const Ham& H, D d, int* occ) {
    int k = 0;
    for (int i = 0; i < H.n_so; ++i)
        if (d >> i & 1) occ[k++] = i;
    return k;
}

// ---- Slater-Condon, exact (unique emission) ---------------------------------------

// DiagCal_bit restated with exact accumulation: E_core + sum_occ h_mm +
// sum_{m != p occ} 1/2 [(mm|pp) - (mp|pm)].
template <class D>
static double diag_exact(const Ham& H, D d) {
    int occ[128];
    int k = occ_list(H, d, occ);
    fx_t acc = fx_from(H.constant);
    for (int a = 0; a < k; ++a) acc += fx_from(H.h[occ[a] * H.n_so + occ[a]]);
    for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b) {
            if (a == b) continue;
            int m = occ[a], p = occ[b];
            acc += fx_from(0.5 * H.g(m, m, p, p));
            acc -= fx_from(0.5 * H.g(m, p, p, m));
        }
    return fx_to(acc);
}

// <D'|H|D> for D' = D with n -> m (m empty, n occupied).
template <class D>
static double single_exact(const Ham& H, D d, int n, int m) {
    int occ[128];
    int k = occ_list(H, d, occ);
    fx_t acc = fx_from(H.h[m * H.n_so + n]);
    for (int a = 0; a < k; ++a) {
        int p = occ[a];
        if (p == n) continue;
        acc += fx_from(H.g(m, n, p, p));
        acc -= fx_from(H.g(m, p, p, n));
    }
    return sgn_between(d, m, n) * fx_to(acc);
}

// <D''|H|D> for D'' = D - {n,q} + {x,y}, n < q occupied, x,y empty.
template <class D>
static double double_exact(const Ham& H, D d, int n, int q, int x, int y) {
    D one = 1;
    D d1 = d - (one << q) + (one << y);
    double s1 = sgn_between(d, q, y) * sgn_between(d1, n, x);
    D d2 = d - (one << q) + (one << x);
    double s2 = sgn_between(d, q, x) * sgn_between(d2, n, y);
    return s1 * H.g(x, n, y, q) + s2 * H.g(y, n, x, q);
}

// Unique connections of d in a fixed class order (diagonal first, then singles and
// doubles by ascending spin-orbital indices).  Calls f(key, H_ij) for each.
template <class D, class F>
static void connections_exact(const Ham& H, D d, F&& f) {
    int occ[128], vir[128];
    int k = 0, v = 0;
    for (int i = 0; i < H.n_so; ++i) {
        if (d >> i & 1) occ[k++] = i; else vir[v++] = i;
    }
    D one = 1;
    f(d, diag_exact(H, d));
    for (int a = 0; a < k; ++a)
        for (int b = 0; b < v; ++b) {
            int n = occ[a], m = vir[b];
            if ((n ^ m) & 1) continue;
            f(d - (one << n) + (one << m), single_exact(H, d, n, m));
        }
    for (int a = 0; a < k; ++a)
        for (int a2 = a + 1; a2 < k; ++a2) {
            int n = occ[a], q = occ[a2];
            for (int b = 0; b < v; ++b)
                for (int b2 = b + 1; b2 < v; ++b2) {
                    int x = vir[b], y = vir[b2];
                    // spin conservation: {x,y} has the spins of {n,q}
                    int sn = (n & 1) + (q & 1), sx = (x & 1) + (y & 1);
                    if (sn != sx) continue;
                    D t = d - (one << n) - (one << q) + (one << x) + (one << y);
                    f(t, double_exact(H, d, n, q, x, y));
                }
        }
}

// ---- Slater-Condon, literal (reference emission) -----------------------------------

static double diag_literal(const Ham& H, det_t reff) {
    det_t one = 1;
    double sum = H.constant;
    for (const auto& e : H.O_diag)
        if (reff & (one << e.n)) sum += e.c;
    for (const auto& e : H.OO_diag) {
        unsigned m = e.m, n = e.n, p = e.p, q = e.q;
        if ((reff & (one << n)) && (reff & (one << q))) {
            if (m == n && p == q) sum += 0.5 * e.c;
            else if (m == q && n == p) sum -= 0.5 * e.c;
        }
    }
    return sum;
}

// OffDiagGen_bit_e order; f(key, value-with-sign) per emitted entry (coefficient 1).
template <class F>
static void offdiag_literal(const Ham& H, det_t reff, F&& f) {
    det_t one = 1;
    int occ[128];
    int ne = occ_list(H, reff, occ);
    for (int i = 0; i < ne; ++i)
        for (int j = i + 1; j < ne; ++j) {
            unsigned idx = occ[i] + occ[j] * H.n_so;
            for (const auto& e : H.OO[idx]) {
                det_t start = reff;
                unsigned m = e.m, n = e.n, p = e.p, q = e.q;
                bool m_empty = !(reff & (one << m));
                bool p_empty = !(reff & (one << p));
                if (m != n && n != p && p != q && m != q && m_empty && p_empty) {
                    double s = sgn_between(start, p, q);
                    start = start - (one << q) + (one << p);
                    s = s * sgn_between(start, m, n);
                    start = start - (one << n) + (one << m);
                    f(start, s * e.c);
                } else if (m != n && n != p && p == q && q != m && m_empty) {
                    double s = sgn_between(start, m, n);
                    start = start - (one << n) + (one << m);
                    f(start, s * e.c);
                } else if (m != n && n == p && p != q && q != m && m_empty) {
                    double s = sgn_between(start, q, m);
                    start = start - (one << q) + (one << m);
                    f(start, -s * e.c);
                }
            }
        }
    for (int i = 0; i < ne; ++i) {
        unsigned idx = occ[i];
        for (const auto& e : H.O[idx]) {
            det_t start = reff;
            unsigned m = e.m, n = e.n;
            if (!(reff & (one << m))) {
                double s = sgn_between(start, m, n);
                start = start - (one << n) + (one << m);
                f(start, s * e.c);
            }
        }
    }
}

// ---- top-K by |c| with the declared tie rule ---------------------------------------

struct KV { det_t k; double v; };

static inline bool amp_before(const KV& a, const KV& b) {   // "a ranks above b"
    double fa = std::fabs(a.v), fb = std::fabs(b.v);
    if (fa != fb) return fa > fb;
    return a.k < b.k;
}

// keep the `cap` highest-ranked entries, return them sorted by key.
static void top_cap_sorted(std::vector<KV>& x, size_t cap) {
    if (x.size() > cap) {
        std::nth_element(x.begin(), x.begin() + cap, x.end(), amp_before);
        x.resize(cap);
    }
    std::sort(x.begin(), x.end(), [](const KV& a, const KV& b) { return a.k < b.k; });
}

struct Cand { det_t k; double v; uint8_t tag; };

}  // namespace

// =====================================================================================
extern "C" {

typedef struct oracle_ham oracle_ham;

// ReadElements (src/functions.cpp:297-395) plus the parameter block of main
// (src/main.cpp:183-223).  Returns NULL on a malformed file.
oracle_ham* oracle_load(const char* path, int n_so, int n_elec) {
    FILE* fp = fopen(path, "r");
    if (!fp) return nullptr;
    Ham* H = new Ham();
    H->n_so = n_so;
    size_t N = n_so;
    H->OO.resize(N * N);
    H->O.resize(N);
    H->G.assign(N * N * N * N, 0.0);
    H->h.assign(N * N, 0.0);
    for (int blk = 0; blk < 7; ++blk) {
        int num_line = 0;
        if (fscanf(fp, "%d", &num_line) != 1) break;
        for (int i = 0; i < num_line; ++i) {
            double c;
            int m, n, p, q;
            if (fscanf(fp, "%lf %d %d %d %d", &c, &m, &n, &p, &q) != 5) { fclose(fp); delete H; return nullptr; }
            if (p == 0 && q == 0) {
                if (m == 0 && n == 0) { H->constant = c; continue; }
                int a = m - 2, b = n - 2;
                if (a < 0 || b < 0 || a >= n_so || b >= n_so) { fclose(fp); delete H; return nullptr; }
                H->h[a * N + b] += c;
                if (m == n) H->O_diag.push_back({(uint8_t)a, (uint8_t)b, c});
                else H->O[b].push_back({(uint8_t)a, (uint8_t)b, c});
            } else {
                int a = m - 2, b = n - 2, e = p - 2, f = q - 2;
                if (a < 0 || b < 0 || e < 0 || f < 0 || a >= n_so || b >= n_so || e >= n_so || f >= n_so) {
                    fclose(fp); delete H; return nullptr;
                }
                H->G[((a * N + b) * N + e) * N + f] += c;
                Ent4 t{(uint8_t)a, (uint8_t)b, (uint8_t)e, (uint8_t)f, c};
                if ((a == b && e == f) || (a == f && e == b)) {
                    H->OO_diag.push_back(t);
                } else if (!(m == n || n == p || p == q || m == q)) {
                    if (b < f) H->OO[b + N * f].push_back(t);
                } else {
                    size_t idx = b < f ? b + N * f : f + N * b;
                    H->OO[idx].push_back(t);
                }
            }
        }
        if (num_line == 1) break;
    }
    fclose(fp);
    int na = n_elec / 2;
    H->n_alpha = na;
    H->nck_list.assign(na, std::vector<uint64_t>(N / 2, 0));
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < (int)N / 2; ++j)
            if (j >= i + 1) H->nck_list[i][j] = binom(j, i + 1);
    H->block_dim = binom(n_so / 2, na);
    H->bucket_bits = (unsigned)std::log2((double)H->block_dim) / 2;
    H->bucket_num = (H->block_dim >> H->bucket_bits) + 1;
    return (oracle_ham*)H;
}

void oracle_free(oracle_ham* h) { delete (Ham*)h; }

double oracle_constant(const oracle_ham* h) { return ((const Ham*)h)->constant; }

double oracle_integral(const oracle_ham* h, int m, int n, int p, int q) {
    return ((const Ham*)h)->g(m, n, p, q);
}
double oracle_one_body(const oracle_ham* h, int m, int n) {
    const Ham* H = (const Ham*)h;
    return H->h[m * H->n_so + n];
}

uint64_t oracle_alpha_rank(const oracle_ham* h, uint64_t d) { return alpha_rank(*(const Ham*)h, d); }

void oracle_bucket_params(const oracle_ham* h, uint64_t* block_dim, unsigned* bucket_bits, uint64_t* bucket_num) {
    const Ham* H = (const Ham*)h;
    *block_dim = H->block_dim;
    *bucket_bits = H->bucket_bits;
    *bucket_num = H->bucket_num;
}

// Connections per determinant, diagonal included: 1 + 2 k v + 2 C(k,2) C(v,2) + (k v)^2.
uint64_t oracle_n_conn(const oracle_ham* hp) {
    const Ham& H = *(const Ham*)hp;
    uint64_t k = (uint64_t)H.n_alpha, v = (uint64_t)(H.n_so / 2) - k;
    return 1 + 2 * k * v + 2 * (k * (k - 1) / 2) * (v * (v - 1) / 2) + (k * v) * (k * v);
}

double oracle_diag(const oracle_ham* h, uint64_t d) { return diag_exact(*(const Ham*)h, d); }
double oracle_diag_literal(const oracle_ham* h, uint64_t d) { return diag_literal(*(const Ham*)h, d); }

// Unique connections (diagonal first).  Returns the count; writes up to cap.
int64_t oracle_connections(const oracle_ham* h, uint64_t d, uint64_t* keys, double* vals, int64_t cap) {
    int64_t n = 0;
    connections_exact(*(const Ham*)h, d, [&](det_t k, double v) {
        if (n < cap) { keys[n] = k; vals[n] = v; }
        ++n;
    });
    return n;
}

// OffDiagGen_bit_e emission (no diagonal).  Returns the count; writes up to cap.
int64_t oracle_offdiag_literal(const oracle_ham* h, uint64_t d, uint64_t* keys, double* vals, int64_t cap) {
    int64_t n = 0;
    offdiag_literal(*(const Ham*)h, d, [&](det_t k, double v) {
        if (n < cap) { keys[n] = k; vals[n] = v; }
        ++n;
    });
    return n;
}

// SpMVMtply_bucket (src/functions.cpp:511-804).
//   out = top_{out_cap}( cut_{|y| > cut}( coeff (H - lambda) top_nd(in) ) ), key-sorted
//   diag = <top_nd(in)| coeff (H - lambda) |top_nd(in)>, before the cut.
// `in` must be key-sorted.  Returns 0, or -1 when a term leaves the fixed-point range.
int oracle_hv(const oracle_ham* hp, const uint64_t* in_k, const double* in_v, int64_t n_in,
              int64_t nd, double lambda, double coeff, double cut, int64_t out_cap,
              uint64_t* out_k, double* out_v, int64_t* out_len, double* diag_out,
              int nthreads, int mode, int64_t* n_candidates) {
    const Ham& H = *(const Ham*)hp;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    std::vector<KV> src(n_in);
    for (int64_t i = 0; i < n_in; ++i) src[i] = {in_k[i], in_v[i]};
    if (n_in > nd) top_cap_sorted(src, (size_t)nd);     // top-Nd of the input, key order
    int64_t ns = (int64_t)src.size();
    if (ns == 0) { *out_len = 0; *diag_out = 0.0; if (n_candidates) *n_candidates = 0; return 0; }

    const uint64_t nb = H.bucket_num;
    const unsigned bb = H.bucket_bits;
    const int T = nthreads;
    // resGlobal[thread + T*bucket] (src/functions.cpp:565-617)
    std::vector<std::vector<Cand>> res((size_t)T * nb);
    int bad = 0;
#pragma omp parallel num_threads(T)
    {
        int tid = omp_get_thread_num();
#pragma omp for schedule(static, 25)
        for (int64_t i = 0; i < ns; ++i) {
            det_t dk = src[i].k;
            double temp = coeff * src[i].v;
            if (mode != 1) {
                bool first = true;
                connections_exact(H, dk, [&](det_t t, double hij) {
                    double val;
                    uint8_t tag = 0;
                    if (first) { val = (hij - lambda) * temp; tag = 1; first = false; }
                    else if (mode == 2) val = (double)((float)hij * (float)temp);
                    else val = hij * temp;
                    if (!fx_ok(val)) bad = 1;
                    res[tid + (size_t)T * (alpha_rank(H, t) >> bb)].push_back({t, val, tag});
                });
            } else {
                offdiag_literal(H, dk, [&](det_t t, double v) {
                    res[tid + (size_t)T * (alpha_rank(H, t) >> bb)].push_back({t, temp * v, 0});
                });
                double dv = (diag_literal(H, dk) - lambda) * temp;
                res[tid + (size_t)T * (alpha_rank(H, dk) >> bb)].push_back({dk, dv, 1});
            }
        }
    }
    if (bad) return -1;
    // per bucket: concatenate, sort, merge, dot with the input (src/functions.cpp:646-731)
    std::vector<std::vector<KV>> merged(nb);
    std::vector<fx_t> diag_fx(nb, 0);
    std::vector<double> diag_lit(nb, 0.0);
    int64_t total_cand = 0;
    for (auto& r : res) total_cand += (int64_t)r.size();
#pragma omp parallel for schedule(dynamic) num_threads(T)
    for (int64_t b = 0; b < (int64_t)nb; ++b) {
        std::vector<Cand> buf;
        size_t tot = 0;
        for (int t = 0; t < T; ++t) tot += res[t + (size_t)T * b].size();
        buf.reserve(tot);
        for (int t = 0; t < T; ++t) {
            auto& r = res[t + (size_t)T * b];
            buf.insert(buf.end(), r.begin(), r.end());
            std::vector<Cand>().swap(r);
        }
        std::stable_sort(buf.begin(), buf.end(), [](const Cand& a, const Cand& c) { return a.k < c.k; });
        auto& out = merged[b];
        size_t i = 0;
        while (i < buf.size()) {
            size_t j = i;
            bool is_input = false;
            double y;
            if (mode != 1) {
                fx_t acc = 0;
                for (; j < buf.size() && buf[j].k == buf[i].k; ++j) { acc += fx_from(buf[j].v); is_input |= buf[j].tag; }
                y = fx_to(acc);
            } else {
                y = buf[i].v;
                is_input |= buf[i].tag;
                for (j = i + 1; j < buf.size() && buf[j].k == buf[i].k; ++j) { y += buf[j].v; is_input |= buf[j].tag; }
            }
            out.push_back({buf[i].k, y});
            if (is_input) {
                // the input coefficient of this determinant
                auto it = std::lower_bound(src.begin(), src.end(), buf[i].k,
                                           [](const KV& a, det_t k) { return a.k < k; });
                double c = it->v;
                if (mode != 1) diag_fx[b] += fx_from(y * c);
                else diag_lit[b] += y * c;
            }
            i = j;
        }
    }
    double diag;
    if (mode != 1) {
        fx_t acc = 0;
        for (auto v : diag_fx) acc += v;
        diag = fx_to(acc);
    } else {
        diag = 0.0;
        for (auto v : diag_lit) diag += v;
    }
    // cut (strict >) then top out_cap by |c|, key order (src/functions.cpp:741-793)
    std::vector<KV> fin;
    for (auto& m : merged)
        for (auto& e : m)
            if (std::fabs(e.v) > cut) fin.push_back(e);
    top_cap_sorted(fin, (size_t)out_cap);
    for (size_t i = 0; i < fin.size(); ++i) { out_k[i] = fin[i].k; out_v[i] = fin[i].v; }
    *out_len = (int64_t)fin.size();
    *diag_out = diag;
    if (n_candidates) *n_candidates = total_cand;
    return 0;
}

// ---- screened H.v (no reference counterpart) ---------------------------------------
// The screened product keeps a subset of the off-diagonal candidates of the mixed-precision
// product (value (float)H_ij * (float)(coeff c_j)) and every diagonal term, and merges them
// exactly as oracle_hv(mode 2) does.  Step 1 lists the candidates that pass the prefilter
// |value| > (float)eps; the caller decides which to keep; step 2 merges the kept ones.

// Step 1: candidates of top_nd(in) in source (key) order and connections_exact order,
// with their excitation rank (1 single, 2 double).  Returns the count; writes up to cap.
int64_t oracle_screen_candidates(const oracle_ham* hp, const uint64_t* in_k, const double* in_v, int64_t n_in,
                                 int64_t nd, double coeff, double eps, uint64_t* t_out, float* v_out,
                                 uint32_t* cls_out, int64_t cap) {
    const Ham& H = *(const Ham*)hp;
    std::vector<KV> src(n_in);
    for (int64_t i = 0; i < n_in; ++i) src[i] = {in_k[i], in_v[i]};
    if (n_in > nd) top_cap_sorted(src, (size_t)nd);
    const float feps = (float)eps;
    int64_t n = 0;
    for (const KV& s : src) {
        const double temp = coeff * s.v;
        bool first = true;
        connections_exact(H, s.k, [&](det_t t, double hij) {
            if (first) {   // the diagonal: always merged, in FP64
                first = false;
                return;
            }
            const float val = (float)hij * (float)temp;
            if (!(std::fabs(val) > feps)) return;
            if (n < cap) {
                t_out[n] = t;
                v_out[n] = val;
                cls_out[n] = (uint32_t)(__builtin_popcountll(t ^ s.k) / 2);
            }
            ++n;
        });
    }
    return n;
}

// Step 2: y_t = sum of the kept candidates' values for t + (E_tt - lambda) coeff c_t for t in
// top_nd(in), exact; diag = <top_nd(in)|y> (exact, before the cut); out = top_{out_cap} of
// {|y| > cut}, key order.  Returns 0, or -1 when a term leaves the fixed-point range.
int oracle_screen_merge(const oracle_ham* hp, const uint64_t* in_k, const double* in_v, int64_t n_in, int64_t nd,
                        double lambda, double coeff, double cut, int64_t out_cap, const uint64_t* ct, const float* cv,
                        int64_t nc, uint64_t* out_k, double* out_v, int64_t* out_len, double* diag_out) {
    const Ham& H = *(const Ham*)hp;
    std::vector<KV> src(n_in);
    for (int64_t i = 0; i < n_in; ++i) src[i] = {in_k[i], in_v[i]};
    if (n_in > nd) top_cap_sorted(src, (size_t)nd);
    std::vector<Cand> buf;
    buf.reserve((size_t)nc + src.size());
    for (int64_t i = 0; i < nc; ++i) {
        if (!fx_ok((double)cv[i])) return -1;
        buf.push_back({ct[i], (double)cv[i], 0});
    }
    for (const KV& s : src) {
        const double dv = (diag_exact(H, s.k) - lambda) * (coeff * s.v);
        if (!fx_ok(dv)) return -1;
        buf.push_back({s.k, dv, 1});
    }
    std::stable_sort(buf.begin(), buf.end(), [](const Cand& a, const Cand& c) { return a.k < c.k; });
    std::vector<KV> fin;
    fx_t dacc = 0;
    size_t i = 0;
    while (i < buf.size()) {
        size_t j = i;
        fx_t acc = 0;
        bool is_input = false;
        for (; j < buf.size() && buf[j].k == buf[i].k; ++j) {
            acc += fx_from(buf[j].v);
            is_input |= buf[j].tag;
        }
        const double y = fx_to(acc);
        if (is_input) {
            auto it = std::lower_bound(src.begin(), src.end(), buf[i].k, [](const KV& a, det_t k) { return a.k < k; });
            dacc += fx_from(y * it->v);
        }
        if (std::fabs(y) > cut) fin.push_back({buf[i].k, y});
        i = j;
    }
    top_cap_sorted(fin, (size_t)out_cap);
    for (size_t q = 0; q < fin.size(); ++q) {
        out_k[q] = fin[q].k;
        out_v[q] = fin[q].v;
    }
    *out_len = (int64_t)fin.size();
    *diag_out = fx_to(dacc);
    return 0;
}

// ---- sparse-vector operations (src/all.h) ------------------------------------------

// dynVec::dot (src/all.h:317-351), exact accumulation.
double oracle_dot(const uint64_t* ak, const double* av, int64_t na,
                  const uint64_t* bk, const double* bv, int64_t nb) {
    int64_t i = 0, j = 0;
    fx_t acc = 0;
    while (i < na && j < nb) {
        if (ak[i] < bk[j]) ++i;
        else if (ak[i] > bk[j]) ++j;
        else { acc += fx_from(av[i] * bv[j]); ++i; ++j; }
    }
    return fx_to(acc);
}

// dynVec::norm (src/all.h:249-258), exact accumulation.
double oracle_norm(const double* v, int64_t n) {
    fx_t acc = 0;
    for (int64_t i = 0; i < n; ++i) acc += fx_from(v[i] * v[i]);
    return std::sqrt(fx_to(acc));
}

// dynVec::addTwo (src/all.h:407-480): v <- top_cap(merge(v + a*w)), zeros kept, key order.
// out may alias nothing; returns the new length.
int64_t oracle_add_two(const uint64_t* vk, const double* vv, int64_t nv,
                       const uint64_t* wk, const double* wv, int64_t nw, double a,
                       int64_t cap, uint64_t* ok, double* ov) {
    std::vector<KV> t;
    t.reserve(nv + nw);
    int64_t i = 0, j = 0;
    while (i < nv && j < nw) {
        if (vk[i] < wk[j]) { t.push_back({vk[i], vv[i]}); ++i; }
        else if (vk[i] > wk[j]) { t.push_back({wk[j], a * wv[j]}); ++j; }
        else { t.push_back({wk[j], a * wv[j] + vv[i]}); ++i; ++j; }
    }
    for (; i < nv; ++i) t.push_back({vk[i], vv[i]});
    for (; j < nw; ++j) t.push_back({wk[j], a * wv[j]});
    top_cap_sorted(t, (size_t)cap);
    for (size_t k = 0; k < t.size(); ++k) { ok[k] = t[k].k; ov[k] = t[k].v; }
    return (int64_t)t.size();
}

// dynVec::shrinkFrom (src/all.h:281-299): top-cap by |c|, key order.
int64_t oracle_shrink(const uint64_t* sk, const double* sv, int64_t ns, int64_t cap,
                      uint64_t* ok, double* ov) {
    std::vector<KV> t(ns);
    for (int64_t i = 0; i < ns; ++i) t[i] = {sk[i], sv[i]};
    top_cap_sorted(t, (size_t)cap);
    for (size_t k = 0; k < t.size(); ++k) { ok[k] = t[k].k; ov[k] = t[k].v; }
    return (int64_t)t.size();
}

// Restart recombination (src/sparse_solver.cpp:492-517): sum_i c_i K_i merged (exact),
// then top-cap and normalize.  Vectors are given concatenated with offsets.
int64_t oracle_lincomb(const uint64_t* kk, const double* kv, const int64_t* offs, int nvec,
                       const double* c, int64_t cap, uint64_t* ok, double* ov, double* norm_out) {
    std::vector<std::pair<det_t, double>> all;
    for (int i = 0; i < nvec; ++i)
        for (int64_t e = offs[i]; e < offs[i + 1]; ++e) all.push_back({kk[e], c[i] * kv[e]});
    std::stable_sort(all.begin(), all.end(), [](const std::pair<det_t, double>& a, const std::pair<det_t, double>& b) { return a.first < b.first; });
    std::vector<KV> m;
    size_t i = 0;
    while (i < all.size()) {
        fx_t acc = 0;
        size_t j = i;
        for (; j < all.size() && all[j].first == all[i].first; ++j) acc += fx_from(all[j].second);
        m.push_back({all[i].first, fx_to(acc)});
        i = j;
    }
    fx_t nacc = 0;
    for (auto& e : m) nacc += fx_from(e.v * e.v);
    if (norm_out) *norm_out = std::sqrt(fx_to(nacc));
    top_cap_sorted(m, (size_t)cap);
    fx_t tacc = 0;
    for (auto& e : m) tacc += fx_from(e.v * e.v);
    double inv = 1.0 / std::sqrt(fx_to(tacc));
    for (size_t k = 0; k < m.size(); ++k) { ok[k] = m[k].k; ov[k] = m[k].v * inv; }
    return (int64_t)m.size();
}

// calc_energy (src/functions.cpp:1272-1313): <v|H|v> = sum_i c_i (c_i H_ii + sum_{t in v,
// t != i} H_ti c_t), per source as the reference pulls it, with exact sums:
// E_i = round(fx(c_i H_ii) + sum fx(H_ti c_t)), energy = round(sum fx(c_i E_i)).
double oracle_energy(const oracle_ham* h, const uint64_t* k, const double* v, int64_t n, int nthreads) {
    const Ham& H = *(const Ham*)h;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    std::vector<fx_t> part(nthreads, 0);
#pragma omp parallel num_threads(nthreads)
    {
        fx_t mine = 0;
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < n; ++i) {
            const det_t d = k[i];
            const double ci = v[i];
            fx_t e = 0;
            bool first = true;
            connections_exact(H, d, [&](det_t t, double hij) {
                if (first) {   // the diagonal comes first
                    e += fx_from(ci * hij);
                    first = false;
                    return;
                }
                const uint64_t* it = std::lower_bound(k, k + n, t);
                if (it != k + n && *it == t) e += fx_from(hij * v[it - k]);
            });
            mine += fx_from(ci * fx_to(e));
        }
        part[omp_get_thread_num()] = mine;
    }
    fx_t tot = 0;
    for (auto p : part) tot += p;
    return fx_to(tot);
}

// ---- 128-bit determinant keys (more than 64 spin orbitals) -------------------------
// No reference counterpart: the reference stores determinants in one 64-bit word
// (src/all.h:150-194, src/main.cpp:84-85).  These functions extend the exact mode (0)
// and the mixed mode (2) of oracle_hv to keys of two words, lo = spin orbitals 0-63,
// hi = spin orbitals 64-127, ordered as the 128-bit integer hi:lo.  The Slater-Condon
// helpers are the same templates the 64-bit functions use.

// Unique connections of one determinant (diagonal first).  Returns the count; writes up
// to cap.
int64_t oracle_connections_wide(const oracle_ham* h, uint64_t lo, uint64_t hi, uint64_t* out_lo,
                                uint64_t* out_hi, double* vals, int64_t cap) {
    const u128 d = ((u128)hi << 64) | lo;
    int64_t n = 0;
    connections_exact(*(const Ham*)h, d, [&](u128 k, double v) {
        if (n < cap) { out_lo[n] = (uint64_t)k; out_hi[n] = (uint64_t)(k >> 64); vals[n] = v; }
        ++n;
    });
    return n;
}

double oracle_diag_wide(const oracle_ham* h, uint64_t lo, uint64_t hi) {
    return diag_exact(*(const Ham*)h, ((u128)hi << 64) | lo);
}

// oracle_hv (modes 0 and 2) on 128-bit keys; `in` key-sorted.  Returns 0, -1 when a term
// leaves the fixed-point range, -2 for an unsupported mode.
int oracle_hv_wide(const oracle_ham* hp, const uint64_t* in_lo, const uint64_t* in_hi, const double* in_v,
                   int64_t n_in, int64_t nd, double lambda, double coeff, double cut, int64_t out_cap,
                   uint64_t* out_lo, uint64_t* out_hi, double* out_v, int64_t* out_len, double* diag_out,
                   int nthreads, int mode) {
    const Ham& H = *(const Ham*)hp;
    if (mode != 0 && mode != 2) return -2;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    struct KW { u128 k; double v; };
    auto rank_above = [](const KW& a, const KW& b) {
        double fa = std::fabs(a.v), fb = std::fabs(b.v);
        if (fa != fb) return fa > fb;
        return a.k < b.k;
    };
    auto by_key = [](const KW& a, const KW& b) { return a.k < b.k; };
    auto top_cap = [&](std::vector<KW>& x, size_t cap) {
        if (x.size() > cap) {
            std::nth_element(x.begin(), x.begin() + cap, x.end(), rank_above);
            x.resize(cap);
        }
        std::sort(x.begin(), x.end(), by_key);
    };
    std::vector<KW> src(n_in);
    for (int64_t i = 0; i < n_in; ++i) src[i] = {((u128)in_hi[i] << 64) | in_lo[i], in_v[i]};
    if (n_in > nd) top_cap(src, (size_t)nd);
    const int64_t ns = (int64_t)src.size();
    // generation into per-(thread, bucket) lists, bucket = a hash of the target key; then
    // every bucket sorted and merged on its own (the decomposition of oracle_hv)
    struct CW { u128 k; double v; uint8_t tag; };
    const int T = nthreads;
    const size_t nb = (size_t)4 * T;
    auto bucket_of = [nb](u128 k) {
        uint64_t z = (uint64_t)k * 0x9E3779B97F4A7C15ull ^ (uint64_t)(k >> 64) * 0xC2B2AE3D27D4EB4Full;
        z ^= z >> 29;
        return (size_t)(z % nb);
    };
    std::vector<std::vector<CW>> res((size_t)T * nb);
    int bad = 0;
#pragma omp parallel num_threads(T)
    {
        const int tid = omp_get_thread_num();
#pragma omp for schedule(static, 25)
        for (int64_t i = 0; i < ns; ++i) {
            const double temp = coeff * src[i].v;
            bool first = true;
            connections_exact(H, src[i].k, [&](u128 t, double hij) {
                double val;
                uint8_t tag = 0;
                if (first) { val = (hij - lambda) * temp; tag = 1; first = false; }
                else if (mode == 2) val = (double)((float)hij * (float)temp);
                else val = hij * temp;
                if (!fx_ok(val)) bad = 1;
                res[tid + (size_t)T * bucket_of(t)].push_back({t, val, tag});
            });
        }
    }
    if (bad) return -1;
    std::vector<std::vector<KW>> kept(nb);
    std::vector<fx_t> dpart(nb, 0);
#pragma omp parallel for schedule(dynamic) num_threads(T)
    for (int64_t b = 0; b < (int64_t)nb; ++b) {
        std::vector<CW> buf;
        for (int t = 0; t < T; ++t) {
            auto& r = res[t + (size_t)T * b];
            buf.insert(buf.end(), r.begin(), r.end());
            std::vector<CW>().swap(r);
        }
        std::sort(buf.begin(), buf.end(), [](const CW& a, const CW& c) { return a.k < c.k; });
        size_t i = 0;
        while (i < buf.size()) {
            size_t j = i;
            fx_t acc = 0;
            bool is_input = false;
            for (; j < buf.size() && buf[j].k == buf[i].k; ++j) { acc += fx_from(buf[j].v); is_input |= buf[j].tag; }
            const double y = fx_to(acc);
            if (is_input) {
                auto it = std::lower_bound(src.begin(), src.end(), buf[i].k,
                                           [](const KW& a, u128 k) { return a.k < k; });
                dpart[b] += fx_from(y * it->v);
            }
            if (std::fabs(y) > cut) kept[b].push_back({buf[i].k, y});
            i = j;
        }
    }
    fx_t dacc = 0;
    for (auto d : dpart) dacc += d;
    std::vector<KW> fin;
    for (auto& k : kept) fin.insert(fin.end(), k.begin(), k.end());
    top_cap(fin, (size_t)out_cap);
    for (size_t q = 0; q < fin.size(); ++q) {
        out_lo[q] = (uint64_t)fin[q].k;
        out_hi[q] = (uint64_t)(fin[q].k >> 64);
        out_v[q] = fin[q].v;
    }
    *out_len = (int64_t)fin.size();
    *diag_out = fx_to(dacc);
    return 0;
}

}  // extern "C"
