// CPU restatement of the reference's truncated Lanczos driver -- TEST INFRASTRUCTURE.
//
// sparse_krylov_solve_eigenval (src/sparse_solver.cpp:246-548) on top of the oracle's
// H.v and sparse-vector operations (sfci_oracle.cpp).  The Nl x Nl generalized
// eigenproblem that the reference hands to Eigen's GeneralizedSelfAdjointEigenSolver
// (src/sparse_solver.cpp:450-454) is solved by the library's dense solver
// (sfci_amd/csrc/gsaes.h: Cholesky of S, cyclic Jacobi on L^-1 K L^-T), which is host
// math off the GPU path and is checked on its own by residual tests.  Eigenvector sign:
// the first component with |c_i| > 0 is made positive (the reference takes whatever
// sign Eigen returns).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../sfci_amd/csrc/gsaes.h"

extern "C" {
typedef struct oracle_ham oracle_ham;
int oracle_hv(const oracle_ham*, const uint64_t*, const double*, int64_t, int64_t, double, double, double,
              int64_t, uint64_t*, double*, int64_t*, double*, int, int, int64_t*);
double oracle_dot(const uint64_t*, const double*, int64_t, const uint64_t*, const double*, int64_t);
double oracle_norm(const double*, int64_t);
int64_t oracle_add_two(const uint64_t*, const double*, int64_t, const uint64_t*, const double*, int64_t,
                       double, int64_t, uint64_t*, double*);
int64_t oracle_shrink(const uint64_t*, const double*, int64_t, int64_t, uint64_t*, double*);
int64_t oracle_lincomb(const uint64_t*, const double*, const int64_t*, int, const double*, int64_t,
                       uint64_t*, double*, double*);
double oracle_energy(const oracle_ham*, const uint64_t*, const double*, int64_t, int);
}

using sfci_dense::gsaes;

namespace {

struct Vec {
    std::vector<uint64_t> k;
    std::vector<double> v;
    int64_t len = 0;
    int64_t cap = 0;
    explicit Vec(int64_t c = 0) : k(c), v(c), cap(c) {}
    void copy_from(const Vec& s) {
        if (s.len > (int64_t)k.size()) { k.resize(s.len); v.resize(s.len); }
        std::copy(s.k.begin(), s.k.begin() + s.len, k.begin());
        std::copy(s.v.begin(), s.v.begin() + s.len, v.begin());
        len = s.len;
    }
};

}  // namespace

extern "C" {

int oracle_gsaes(int n, const double* K, const double* S, double* evals, double* evecs) {
    std::vector<double> k(K, K + n * n), s(S, S + n * n), e, v;
    int rc = gsaes(n, k, s, e, v);
    if (rc) return rc;
    std::copy(e.begin(), e.end(), evals);
    std::copy(v.begin(), v.end(), evecs);
    return 0;
}

// Least squares of the Krylov linear solver (gsaes.h lstsq); returns the rank.
int oracle_lstsq(int m, int n, const double* A, const double* b, double* x) {
    std::vector<double> a(A, A + (size_t)m * n), bb(b, b + m), xx;
    int r = sfci_dense::lstsq(m, n, a, bb, xx);
    std::copy(xx.begin(), xx.end(), x);
    return r;
}

// sparse_krylov_solve_eigenval (src/sparse_solver.cpp:246-548).
// Start vector: init_det with coefficient 1.  Writes per-restart Ritz values and
// <v|H|v> after truncation (up to max_restart each); returns the number of restarts run
// or a negative error.  Final vector written to out_k/out_v (capacity nd).
int oracle_lanczos(const oracle_ham* H, uint64_t init_det, int64_t nd, int nl, int max_restart,
                   int64_t hpsi_cap, double cut, int nthreads, double e_prev_init,
                   double* ritz, double* e_trunc, double* e_final,
                   uint64_t* out_k, double* out_v, int64_t* out_len) {
    const double deltaE = 1e-5;
    double previousE = e_prev_init, nearestE = -1000000.0;
    std::vector<Vec> K(nl, Vec(nd));
    Vec v(nd);
    v.k[0] = init_det; v.v[0] = 1.0; v.len = 1;
    std::vector<double> KM(nl * nl, 0.0), S(nl * nl, 0.0);
    int res = 0;
    for (res = 0; res < max_restart; ++res) {
        K[0].copy_from(v);
        Vec u(hpsi_cap), temp(hpsi_cap), temp_u(hpsi_cap), v_next(nd);
        double alpha;
        int rc = oracle_hv(H, v.k.data(), v.v.data(), v.len, nd, 0.0, 1.0, cut, u.cap,
                           u.k.data(), u.v.data(), &u.len, &alpha, nthreads, 0, nullptr);
        if (rc) return -2;
        KM[0] = alpha;
        double beta;
        for (int i = 1; i < nl; ++i) {
            temp.copy_from(u);
            temp.len = oracle_add_two(u.k.data(), u.v.data(), u.len, v.k.data(), v.v.data(), v.len,
                                      -alpha, temp.cap, temp.k.data(), temp.v.data());
            v_next.len = oracle_shrink(temp.k.data(), temp.v.data(), temp.len, v_next.cap,
                                       v_next.k.data(), v_next.v.data());
            double g = oracle_norm(v_next.v.data(), v_next.len);
            beta = g;
            double inv = 1.0 / g;
            for (int64_t t = 0; t < v_next.len; ++t) v_next.v[t] *= inv;
            rc = oracle_hv(H, v_next.k.data(), v_next.v.data(), v_next.len, nd, 0.0, 1.0, cut, temp_u.cap,
                           temp_u.k.data(), temp_u.v.data(), &temp_u.len, &alpha, nthreads, 0, nullptr);
            if (rc) return -2;
            KM[i * nl + i] = alpha;
            for (int j = 0; j < i; ++j) {
                double d = oracle_dot(K[j].k.data(), K[j].v.data(), K[j].len,
                                      temp_u.k.data(), temp_u.v.data(), temp_u.len);
                KM[j * nl + i] = KM[i * nl + j] = d;
            }
            Vec t2(hpsi_cap);
            t2.len = oracle_add_two(temp_u.k.data(), temp_u.v.data(), temp_u.len, v.k.data(), v.v.data(), v.len,
                                    -beta, t2.cap, t2.k.data(), t2.v.data());
            u.copy_from(t2);
            v.copy_from(v_next);
            K[i].copy_from(v);
        }
        for (int i = 0; i < nl; ++i)
            for (int j = i; j < nl; ++j) {
                double d = oracle_dot(K[i].k.data(), K[i].v.data(), K[i].len, K[j].k.data(), K[j].v.data(), K[j].len);
                S[i * nl + j] = S[j * nl + i] = d;
            }
        std::vector<double> ev, evec;
        if (gsaes(nl, KM, S, ev, evec)) return -3;
        int nearest = 0;
        double maxGap = 0;
        for (int i = 0; i < nl; ++i) {
            double gap = std::fabs(nearestE - ev[i]);
            if (i == 0 || gap < maxGap) { maxGap = gap; nearest = i; }
        }
        nearestE = ev[nearest];
        std::vector<double> c(nl);
        for (int i = 0; i < nl; ++i) c[i] = evec[i * nl + nearest];
        std::vector<int64_t> offs(nl + 1, 0);
        for (int i = 0; i < nl; ++i) offs[i + 1] = offs[i] + K[i].len;
        std::vector<uint64_t> ck(offs[nl]);
        std::vector<double> cv(offs[nl]);
        for (int i = 0; i < nl; ++i) {
            std::copy(K[i].k.begin(), K[i].k.begin() + K[i].len, ck.begin() + offs[i]);
            std::copy(K[i].v.begin(), K[i].v.begin() + K[i].len, cv.begin() + offs[i]);
        }
        v.len = oracle_lincomb(ck.data(), cv.data(), offs.data(), nl, c.data(), v.cap, v.k.data(), v.v.data(), nullptr);
        double et = oracle_energy(H, v.k.data(), v.v.data(), v.len, nthreads);
        ritz[res] = ev[nearest];
        e_trunc[res] = et;
        std::fill(KM.begin(), KM.end(), 0.0);
        for (auto& kk : K) kk.len = 0;
        if (previousE - ev[nearest] < deltaE) { ++res; break; }
        previousE = ev[nearest];
    }
    *e_final = oracle_energy(H, v.k.data(), v.v.data(), v.len, nthreads);
    for (int64_t t = 0; t < v.len; ++t) { out_k[t] = v.k[t]; out_v[t] = v.v[t]; }
    *out_len = v.len;
    return res;
}

}  // extern "C"
