// CPU restatement of the reference's linear-solver path -- TEST INFRASTRUCTURE.
//
//   spmspvm_bucket_linear        src/functions.cpp:858-1094
//   spmspvm_bucket_res_norm      src/functions.cpp:1096-1259
//   dynVec::addThree             src/all.h:482-608
//   add_bucketed_vec             src/functions.cpp:807-856
//   sparse_krylov_solve_linear   src/sparse_solver.cpp:561-731
//   sparse_solve_SI              src/sparse_solver.cpp:733-772
//   sparse_main                  src/sparse_solver.cpp:774-809
//   minres_sp                    src/minres.cpp:73-225
//   calc_CISD                    src/sparse_solver.cpp:15-114
// on top of the oracle's H.v and sparse-vector operations (sfci_oracle.cpp).  Every sum
// over duplicate keys is exact (128-bit fixed point), as in the rest of the oracle; the
// reference's bucket-order FP64 sums differ from these in the last bits only.  The
// (m+1) x m least-squares problem the reference hands to Eigen's colPivHouseholderQr
// (src/sparse_solver.cpp:672) and the CISD eigenproblem (Eigen's SelfAdjointEigenSolver,
// src/sparse_solver.cpp:98-100) use the library's dense host code (gsaes.h), checked on
// its own by residual tests.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../sfci_amd/csrc/gsaes.h"
#include "fixacc.h"

extern "C" {
typedef struct oracle_ham oracle_ham;
int oracle_hv(const oracle_ham*, const uint64_t*, const double*, int64_t, int64_t, double, double, double,
              int64_t, uint64_t*, double*, int64_t*, double*, int, int, int64_t*);
double oracle_dot(const uint64_t*, const double*, int64_t, const uint64_t*, const double*, int64_t);
double oracle_norm(const double*, int64_t);
int64_t oracle_add_two(const uint64_t*, const double*, int64_t, const uint64_t*, const double*, int64_t,
                       double, int64_t, uint64_t*, double*);
int64_t oracle_shrink(const uint64_t*, const double*, int64_t, int64_t, uint64_t*, double*);
double oracle_energy(const oracle_ham*, const uint64_t*, const double*, int64_t, int);
int64_t oracle_connections(const oracle_ham*, uint64_t, uint64_t*, double*, int64_t);
uint64_t oracle_n_conn(const oracle_ham*);
}

namespace {

using oracle::fx_from;
using oracle::fx_t;
using oracle::fx_to;

struct V {
    std::vector<uint64_t> k;
    std::vector<double> v;
    int64_t len = 0;
    int64_t cap = 0;
    explicit V(int64_t c = 0) : k(std::max<int64_t>(c, 1)), v(std::max<int64_t>(c, 1)), cap(c) {}
    void copy_from(const V& s) {
        if (s.len > (int64_t)k.size()) { k.resize(s.len); v.resize(s.len); }
        std::copy(s.k.begin(), s.k.begin() + s.len, k.begin());
        std::copy(s.v.begin(), s.v.begin() + s.len, v.begin());
        len = s.len;
    }
    void set(const uint64_t* kk, const double* vv, int64_t n) {
        if (n > (int64_t)k.size()) { k.resize(n); v.resize(n); }
        std::copy(kk, kk + n, k.begin());
        std::copy(vv, vv + n, v.begin());
        len = n;
    }
};

struct KVp {
    uint64_t k;
    double v;
};

// top `cap` by (|c| desc, key asc), returned in key order
void top_cap(std::vector<KVp>& x, size_t cap) {
    if (x.size() > cap) {
        std::nth_element(x.begin(), x.begin() + cap, x.end(), [](const KVp& a, const KVp& b) {
            double fa = std::fabs(a.v), fb = std::fabs(b.v);
            return fa > fb || (fa == fb && a.k < b.k);
        });
        x.resize(cap);
    }
    std::sort(x.begin(), x.end(), [](const KVp& a, const KVp& b) { return a.k < b.k; });
}

// dst = top_{dst.cap}({|c| > cut} of the key-sorted src)
void cut_cap(const V& src, double cut, V& dst) {
    std::vector<KVp> t;
    for (int64_t i = 0; i < src.len; ++i)
        if (std::fabs(src.v[i]) > cut) t.push_back({src.k[i], src.v[i]});
    top_cap(t, (size_t)dst.cap);
    dst.len = (int64_t)t.size();
    if (dst.len > (int64_t)dst.k.size()) { dst.k.resize(dst.len); dst.v.resize(dst.len); }
    for (size_t i = 0; i < t.size(); ++i) { dst.k[i] = t[i].k; dst.v[i] = t[i].v; }
}

// (H - lambda) x with every distinct target kept (cut -1 keeps exact zeros), key order
int hv_all(const oracle_ham* H, const V& x, double lambda, V& w) {
    const int64_t bound = std::max<int64_t>(1, x.len * (int64_t)oracle_n_conn(H));
    w = V(bound);
    double d = 0.0;
    if (x.len == 0) { w.len = 0; return 0; }
    return oracle_hv(H, x.k.data(), x.v.data(), x.len, x.len, lambda, 1.0, -1.0, bound, w.k.data(), w.v.data(),
                     &w.len, &d, 0, 0, nullptr);
}

// a <- a + s b with no truncation (addTwo with room for every key)
void add_into(V& a, const V& b, double s) {
    V t(a.len + b.len);
    t.len = oracle_add_two(a.k.data(), a.v.data(), a.len, b.k.data(), b.v.data(), b.len, s, t.cap, t.k.data(),
                           t.v.data());
    a = t;
}

void add_capped(V& a, const V& b, double s) {
    V t(a.len + b.len);
    t.len = oracle_add_two(a.k.data(), a.v.data(), a.len, b.k.data(), b.v.data(), b.len, s, a.cap, t.k.data(),
                           t.v.data());
    a.copy_from(t);
}

double norm(const V& a) { return oracle_norm(a.v.data(), a.len); }
double dot(const V& a, const V& b) { return oracle_dot(a.k.data(), a.v.data(), a.len, b.k.data(), b.v.data(), b.len); }
void scale(V& a, double s) {
    for (int64_t i = 0; i < a.len; ++i) a.v[i] *= s;
}
void normalize(V& a) { scale(a, 1.0 / norm(a)); }

// sum_i c_i vecs[i], exact merge, top cap (no normalization)
void lincomb(const std::vector<const V*>& vs, const double* c, V& out) {
    std::vector<std::pair<uint64_t, double>> all;
    for (size_t i = 0; i < vs.size(); ++i)
        for (int64_t e = 0; e < vs[i]->len; ++e) all.push_back({vs[i]->k[e], c[i] * vs[i]->v[e]});
    std::stable_sort(all.begin(), all.end(),
                     [](const std::pair<uint64_t, double>& a, const std::pair<uint64_t, double>& b) {
                         return a.first < b.first;
                     });
    std::vector<KVp> m;
    size_t i = 0;
    while (i < all.size()) {
        fx_t acc = 0;
        size_t j = i;
        for (; j < all.size() && all[j].first == all[i].first; ++j) acc += fx_from(all[j].second);
        m.push_back({all[i].first, fx_to(acc)});
        i = j;
    }
    top_cap(m, (size_t)out.cap);
    out.len = (int64_t)m.size();
    if (out.len > (int64_t)out.k.size()) { out.k.resize(out.len); out.v.resize(out.len); }
    for (size_t t = 0; t < m.size(); ++t) { out.k[t] = m[t].k; out.v[t] = m[t].v; }
}

int hv_linear(const oracle_ham* H, const V& in, V& out, double lambda, const std::vector<const V*>& basis,
              double cut, double* hij) {
    const int n = (int)basis.size();
    for (int j = 0; j < n; ++j) hij[j] = 0.0;
    out.len = 0;
    if (in.len == 0) return 0;
    V w;
    if (hv_all(H, in, lambda, w)) return -1;
    for (int j = 0; j < n; ++j) hij[j] = dot(w, *basis[j]);
    add_into(w, *basis[n - 1], -hij[n - 1]);
    if (n > 1) add_into(w, *basis[n - 2], -hij[n - 2]);
    cut_cap(w, cut, out);
    return 0;
}

int hv_residual(const oracle_ham* H, const V& x, V& out, double lambda, const V& b, double cut, double* res) {
    V w;
    if (hv_all(H, x, lambda, w)) return -1;
    add_into(w, b, -1.0);
    *res = norm(w);
    cut_cap(w, cut, out);
    return 0;
}

void add_three(V& out, const V& x, const V& y, const V& z, double a, double b, double c) {
    V t(out.cap);
    const double co[3] = {a, b, c};
    lincomb({&x, &y, &z}, co, t);
    out.len = 0;
    for (int64_t i = 0; i < t.len; ++i)
        if (std::fabs(t.v[i]) > 1e-14) {
            out.k[out.len] = t.k[i];
            out.v[out.len] = t.v[i];
            ++out.len;
        }
}

int linear_solve(const oracle_ham* H, const V& b, double target_e, int64_t nd, int m, int restarts, double cut,
                 V& x_out, double* residual) {
    std::vector<V> K(m + 1, V(nd));
    V w(nd), r0(nd), x_m(nd), result((int64_t)m * nd);
    if (b.len > nd) return -4;
    r0.copy_from(b);
    const int n1 = m + 1;
    std::vector<double> hess(n1 * n1, 0.0), S(n1 * n1, 0.0), hij(n1, 0.0);
    for (int res = 0; res < restarts; ++res) {
        double beta = norm(r0);
        K[0].copy_from(r0);
        normalize(K[0]);
        for (int i = 0; i <= m; ++i) {
            std::vector<const V*> basis;
            for (int j = 0; j <= i; ++j) basis.push_back(&K[j]);
            if (hv_linear(H, K[i], w, target_e, basis, cut, hij.data())) return -1;
            for (int j = i; j >= 0; --j) {
                hess[j * n1 + i] = hess[i * n1 + j] = hij[j];
                double d = dot(K[j], K[i]);
                S[i * n1 + j] = S[j * n1 + i] = d;
            }
            normalize(w);
            if (i < m) K[i + 1].copy_from(w);
        }
        std::vector<double> A(n1 * m, 0.0), e1(n1), y;
        for (int r = 0; r < n1; ++r) {
            for (int c = 0; c < m; ++c) {
                double s = 0.0;
                for (int k = 0; k < n1; ++k) s += S[r * n1 + k] * hess[k * n1 + c];
                A[r * m + c] = s;
            }
            e1[r] = S[r * n1] * beta;
        }
        sfci_dense::lstsq(n1, m, A, e1, y);
        std::vector<const V*> vs;
        for (int j = 0; j < m; ++j) vs.push_back(&K[j]);
        lincomb(vs, y.data(), result);
        add_capped(x_m, result, 1.0);
        double rn = 0.0;
        if (hv_residual(H, x_m, r0, target_e, b, cut, &rn)) return -1;
        scale(r0, -1.0);
        residual[res] = rn;
    }
    x_out.copy_from(x_m);
    return 0;
}

int cisd(const oracle_ham* H, uint64_t ref, int state_idx, V& out, double* eval) {
    const int64_t nc = (int64_t)oracle_n_conn(H);
    std::vector<uint64_t> ck(nc);
    std::vector<double> cv(nc);
    oracle_connections(H, ref, ck.data(), cv.data(), nc);
    std::vector<uint64_t> blk;
    for (int64_t i = 0; i < nc; ++i)
        if (ck[i] == ref || cv[i] != 0.0) blk.push_back(ck[i]);
    std::sort(blk.begin(), blk.end());
    const int n = (int)blk.size();
    if (state_idx < 0 || state_idx >= n) return -1;
    std::vector<double> A((size_t)n * n, 0.0), I((size_t)n * n, 0.0), ev, evec;
    for (int i = 0; i < n; ++i) {
        I[(size_t)i * n + i] = 1.0;
        oracle_connections(H, blk[i], ck.data(), cv.data(), nc);
        for (int64_t e = 0; e < nc; ++e) {
            auto it = std::lower_bound(blk.begin(), blk.end(), ck[e]);
            if (it != blk.end() && *it == ck[e]) A[(size_t)i * n + (it - blk.begin())] = cv[e];
        }
    }
    if (sfci_dense::gsaes(n, A, I, ev, evec)) return -3;
    std::vector<double> x(n);
    for (int i = 0; i < n; ++i) x[i] = evec[(size_t)i * n + state_idx];
    int imax = 0;
    for (int i = 1; i < n; ++i)
        if (std::fabs(x[i]) > std::fabs(x[imax])) imax = i;
    if (x[imax] < 0)
        for (double& t : x) t = -t;
    out = V(n);
    out.set(blk.data(), x.data(), n);
    *eval = ev[state_idx];
    return 0;
}

}  // namespace

extern "C" {

int oracle_hv_linear(const oracle_ham* H, const uint64_t* ik, const double* iv, int64_t n_in, double lambda,
                     const uint64_t* bk, const double* bv, const int64_t* boffs, int n_basis, double cut,
                     int64_t out_cap, uint64_t* ok, double* ov, int64_t* out_len, double* hij) {
    V in;
    in.set(ik, iv, n_in);
    std::vector<V> bs(n_basis);
    std::vector<const V*> bp;
    for (int j = 0; j < n_basis; ++j) {
        bs[j].set(bk + boffs[j], bv + boffs[j], boffs[j + 1] - boffs[j]);
    }
    for (int j = 0; j < n_basis; ++j) bp.push_back(&bs[j]);
    V out(out_cap);
    if (hv_linear(H, in, out, lambda, bp, cut, hij)) return -1;
    std::copy(out.k.begin(), out.k.begin() + out.len, ok);
    std::copy(out.v.begin(), out.v.begin() + out.len, ov);
    *out_len = out.len;
    return 0;
}

int oracle_hv_residual(const oracle_ham* H, const uint64_t* xk, const double* xv, int64_t nx, const uint64_t* bk,
                       const double* bv, int64_t nb, double lambda, double cut, int64_t out_cap, uint64_t* ok,
                       double* ov, int64_t* out_len, double* res) {
    V x, b, out(out_cap);
    x.set(xk, xv, nx);
    b.set(bk, bv, nb);
    if (hv_residual(H, x, out, lambda, b, cut, res)) return -1;
    std::copy(out.k.begin(), out.k.begin() + out.len, ok);
    std::copy(out.v.begin(), out.v.begin() + out.len, ov);
    *out_len = out.len;
    return 0;
}

int64_t oracle_add_three(const uint64_t* xk, const double* xv, int64_t nx, const uint64_t* yk, const double* yv,
                         int64_t ny, const uint64_t* zk, const double* zv, int64_t nz, double a, double b, double c,
                         int64_t cap, uint64_t* ok, double* ov) {
    V x, y, z, out(cap);
    x.set(xk, xv, nx);
    y.set(yk, yv, ny);
    z.set(zk, zv, nz);
    add_three(out, x, y, z, a, b, c);
    std::copy(out.k.begin(), out.k.begin() + out.len, ok);
    std::copy(out.v.begin(), out.v.begin() + out.len, ov);
    return out.len;
}

int oracle_linear_solve(const oracle_ham* H, const uint64_t* bk, const double* bv, int64_t nb, double target_e,
                        int64_t nd, int m, int restarts, double cut, uint64_t* xk, double* xv, int64_t* x_len,
                        double* residual) {
    V b, x(nd);
    b.set(bk, bv, nb);
    int rc = linear_solve(H, b, target_e, nd, m, restarts, cut, x, residual);
    if (rc) return rc;
    std::copy(x.k.begin(), x.k.begin() + x.len, xk);
    std::copy(x.v.begin(), x.v.begin() + x.len, xv);
    *x_len = x.len;
    return 0;
}

// sparse_solve_SI: v = x/|x|, energy <v|H|v>
int oracle_shift_invert(const oracle_ham* H, const uint64_t* bk, const double* bv, int64_t nb, double target_e,
                        int64_t nd, int m, int restarts, double cut, uint64_t* vk, double* vv, int64_t* v_len,
                        double* energy, double* residual) {
    V b, x(nd);
    b.set(bk, bv, nb);
    int rc = linear_solve(H, b, target_e, nd, m, restarts, cut, x, residual);
    if (rc) return rc;
    normalize(x);
    *energy = oracle_energy(H, x.k.data(), x.v.data(), x.len, 0);
    std::copy(x.k.begin(), x.k.begin() + x.len, vk);
    std::copy(x.v.begin(), x.v.begin() + x.len, vv);
    *v_len = x.len;
    return 0;
}

int oracle_cisd(const oracle_ham* H, uint64_t ref, int state_idx, uint64_t* ok, double* ov, int64_t cap,
                int64_t* n_out, double* eval) {
    V out;
    int rc = cisd(H, ref, state_idx, out, eval);
    if (rc) return rc;
    *n_out = out.len;
    if (out.len > cap) return -4;
    std::copy(out.k.begin(), out.k.begin() + out.len, ok);
    std::copy(out.v.begin(), out.v.begin() + out.len, ov);
    return 0;
}

// sparse_main: CISD state 2, normalized, then one shift-invert step
int oracle_excited(const oracle_ham* H, uint64_t init_det, double target_e, int64_t nd, int m, int restarts,
                   double cut, uint64_t* vk, double* vv, int64_t* v_len, double* energy, double* e_cisd) {
    V b0;
    int rc = cisd(H, init_det, 2, b0, e_cisd);
    if (rc) return rc;
    if (b0.len > nd) return -4;
    normalize(b0);
    std::vector<double> res(restarts);
    return oracle_shift_invert(H, b0.k.data(), b0.v.data(), b0.len, target_e, nd, m, restarts, cut, vk, vv, v_len,
                               energy, res.data());
}

// minres_sp (src/minres.cpp:73-225)
int oracle_minres(const oracle_ham* H, const uint64_t* rk, const double* rv, int64_t nr, const uint64_t* gk,
                  const double* gv, int64_t ng, int64_t nd, int iters, double tol, double lambda, double cut,
                  uint64_t* xk, double* xv, int64_t* x_len, int* iters_done, double* res2_out) {
    V rhs, x(nd), v_old(nd), v(nd), v_new(nd), w(nd), w_new(nd), p_oold(nd), p_old(nd), p(nd);
    rhs.set(rk, rv, nr);
    if (ng > nd) return -4;
    x.set(gk, gv, ng);
    double diag = 0.0;
    auto hv = [&](const V& in, V& out, double coeff) {
        return oracle_hv(H, in.k.data(), in.v.data(), in.len, nd, lambda, coeff, cut, out.cap, out.k.data(),
                         out.v.data(), &out.len, &diag, 0, 0, nullptr);
    };
    if (hv(x, v_new, -1.0)) return -1;
    add_capped(v_new, rhs, 1.0);
    double nrm = norm(v_new);
    double residualNorm2 = nrm * nrm;
    w_new.copy_from(v_new);
    double beta_new2 = dot(v_new, w_new);
    double beta_new = std::sqrt(beta_new2);
    const double beta_one = beta_new;
    scale(v_new, 1.0 / beta_new);
    scale(w_new, 1.0 / beta_new);
    double c = 1.0, c_old = 1.0, s = 0.0, s_old = 0.0, eta = 1.0;
    int it = 0;
    while (it < iters) {
        const double beta = beta_new;
        v_old.copy_from(v);
        v.copy_from(v_new);
        w.copy_from(w_new);
        if (hv(w, v_new, 1.0)) return -1;
        add_capped(v_new, v_old, -beta);
        double alpha = dot(v_new, w);
        add_capped(v_new, v, -alpha);
        w_new.copy_from(v_new);
        beta_new2 = dot(v_new, w_new);
        beta_new = std::sqrt(beta_new2);
        scale(v_new, 1.0 / beta_new);
        scale(w_new, 1.0 / beta_new);
        const double r2 = s * alpha + c * c_old * beta;
        const double r3 = s_old * beta;
        const double r1_hat = c * alpha - c_old * s * beta;
        const double r1 = std::sqrt(r1_hat * r1_hat + beta_new * beta_new);
        c_old = c;
        s_old = s;
        c = r1_hat / r1;
        s = beta_new / r1;
        p_oold.copy_from(p_old);
        p_old.copy_from(p);
        add_three(p, w, p_old, p_oold, 1.0 / r1, -r2 / r1, -r3 / r1);
        add_capped(x, p, beta_one * c * eta);
        residualNorm2 *= s * s;
        if (residualNorm2 < tol) break;
        eta = -s * eta;
        ++it;
    }
    std::copy(x.k.begin(), x.k.begin() + x.len, xk);
    std::copy(x.v.begin(), x.v.begin() + x.len, xv);
    *x_len = x.len;
    if (iters_done) *iters_done = it;
    if (res2_out) *res2_out = residualNorm2;
    return 0;
}

}  // extern "C"
