// Linear-solver drivers over the library's C-ABI, host C++ (every vector stays on the GPU;
// the host holds only the (m+1) x (m+1) projected matrices and the MINRES scalars).
//   sfci_linear_solve   sparse_krylov_solve_linear  src/sparse_solver.cpp:561-731
//   sfci_shift_invert   sparse_solve_SI             src/sparse_solver.cpp:733-772
//   sfci_excited        sparse_main                 src/sparse_solver.cpp:774-809
//   sfci_minres         minres_sp                   src/minres.cpp:73-225
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/sfci_gpu.h"
#include "gsaes.h"

namespace {

struct VecGuard {
    std::vector<sfci_vec*> v;
    ~VecGuard() {
        for (auto* p : v) sfci_vec_destroy(p);
    }
};

#define CK(x)                           \
    do {                                \
        int rc_ = (x);                  \
        if (rc_ != SFCI_OK) return rc_; \
    } while (0)

int normalize(sfci_ctx* ctx, sfci_vec* v) {
    double n = 0.0;
    CK(sfci_norm(ctx, v, &n));
    return sfci_scale(ctx, v, 1.0 / n);
}

}  // namespace

extern "C" int sfci_linear_solve(sfci_ctx* ctx, const sfci_vec* b, double target_e, const sfci_linear_params_t* p,
                                 sfci_vec* x_out, double* residual) {
    if (!ctx || !b || !p || !x_out || !residual || p->nd < 1 || p->num_iter < 1 || p->num_restart < 1)
        return SFCI_E_ARG;
    const int m = p->num_iter;
    const int64_t nd = p->nd;
    VecGuard g;
    auto mk = [&](int64_t c, sfci_vec** out) {
        int rc = sfci_vec_create(ctx, c, out);
        if (rc == SFCI_OK) g.v.push_back(*out);
        return rc;
    };
    std::vector<sfci_vec*> K(m + 1);
    for (int i = 0; i <= m; ++i) CK(mk(nd, &K[i]));
    sfci_vec *w, *r0, *x_m, *result;
    CK(mk(nd, &w));
    CK(mk(nd, &r0));
    CK(mk(nd, &x_m));
    // x_m = x_0 + V_m y holds up to m * nd distinct determinants before truncation
    CK(mk((int64_t)m * nd, &result));
    if (sfci_vec_len(b) > nd) return SFCI_E_CAPACITY;
    CK(sfci_vec_copy(r0, b));
    const int n1 = m + 1;
    std::vector<double> hess(n1 * n1, 0.0), S(n1 * n1, 0.0), hij(n1, 0.0);
    for (int res = 0; res < p->num_restart; ++res) {
        double beta = 0.0;
        CK(sfci_norm(ctx, r0, &beta));
        CK(sfci_vec_copy(K[0], r0));
        CK(normalize(ctx, K[0]));
        if (p->verbose) printf("initial residue: %.12g\n", beta);
        for (int i = 0; i <= m; ++i) {
            CK(sfci_hv_linear(ctx, K[i], w, target_e, K.data(), i + 1, hij.data()));
            for (int j = i; j >= 0; --j) {
                hess[j * n1 + i] = hess[i * n1 + j] = hij[j];
                double d = 0.0;
                CK(sfci_dot(ctx, K[j], K[i], &d));
                S[i * n1 + j] = S[j * n1 + i] = d;
            }
            CK(normalize(ctx, w));
            if (i < m) CK(sfci_vec_copy(K[i + 1], w));
        }
        // least squares  (S Hess[:, :m]) y = beta S[:, 0]
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
        // x_m = top_Nd(x_m + V_m y)   (add_bucketed_vec, src/functions.cpp:807-856)
        CK(sfci_lincomb(ctx, K.data(), y.data(), m, result));
        CK(sfci_add_two(ctx, x_m, result, 1.0));
        double rn = 0.0;
        CK(sfci_hv_residual(ctx, x_m, r0, target_e, b, &rn));
        CK(sfci_scale(ctx, r0, -1.0));
        residual[res] = rn;
        if (p->verbose) {
            printf("length of final result: %lld\n", (long long)sfci_vec_len(result));
            printf("residue after truncation: %.12g\n", rn);
            printf("*************************restart %d ends*************************\n", res);
            fflush(stdout);
        }
    }
    CK(sfci_vec_copy(x_out, x_m));
    return sfci_sync(ctx);
}

extern "C" int sfci_shift_invert(sfci_ctx* ctx, const sfci_vec* b, double target_e, const sfci_linear_params_t* p,
                                 sfci_vec* v_out, double* energy, double* residual) {
    if (!ctx || !b || !p || !v_out || !energy || !residual) return SFCI_E_ARG;
    // the reference's loop runs one step (for i < 1); its convergence test never repeats it
    CK(sfci_linear_solve(ctx, b, target_e, p, v_out, residual));
    CK(normalize(ctx, v_out));
    CK(sfci_energy(ctx, v_out, energy));
    if (p->verbose) printf("energy after 0th SI:%.12f\n", *energy);
    return SFCI_OK;
}

extern "C" int sfci_excited(sfci_ctx* ctx, uint64_t init_det, double target_e, const sfci_linear_params_t* p,
                            sfci_vec* v_out, double* energy, double* e_cisd) {
    if (!ctx || !p || !v_out || !energy || !e_cisd || p->num_restart < 1) return SFCI_E_ARG;
    VecGuard g;
    sfci_vec* b0;
    int rc = sfci_vec_create(ctx, p->nd, &b0);
    if (rc) return rc;
    g.v.push_back(b0);
    CK(sfci_cisd(ctx, init_det, 2, b0, e_cisd, nullptr));
    CK(normalize(ctx, b0));
    std::vector<double> res(p->num_restart);
    return sfci_shift_invert(ctx, b0, target_e, p, v_out, energy, res.data());
}

extern "C" int sfci_minres(sfci_ctx* ctx, const sfci_vec* rhs, const sfci_vec* guess, sfci_vec* result, int64_t nd,
                           int iters, double tol, double lambda, int* iters_done, double* res2_out) {
    if (!ctx || !rhs || !guess || !result || nd < 1 || iters < 0) return SFCI_E_ARG;
    VecGuard g;
    auto mk = [&](sfci_vec** out) {
        int rc = sfci_vec_create(ctx, nd, out);
        if (rc == SFCI_OK) g.v.push_back(*out);
        return rc;
    };
    sfci_vec *x, *v_old, *v, *v_new, *w, *w_new, *p_oold, *p_old, *pv;
    CK(mk(&x));
    CK(mk(&v_old));
    CK(mk(&v));
    CK(mk(&v_new));
    CK(mk(&w));
    CK(mk(&w_new));
    CK(mk(&p_oold));
    CK(mk(&p_old));
    CK(mk(&pv));
    if (sfci_vec_len(guess) > nd) return SFCI_E_CAPACITY;
    CK(sfci_vec_copy(x, guess));
    double diag = 0.0;
    // v_new = rhs - A x
    CK(sfci_hv(ctx, x, v_new, nd, lambda, -1.0, &diag));
    CK(sfci_add_two(ctx, v_new, rhs, 1.0));
    double nrm = 0.0;
    CK(sfci_norm(ctx, v_new, &nrm));
    double residualNorm2 = nrm * nrm;
    CK(sfci_vec_copy(w_new, v_new));
    double beta_new2 = 0.0;
    CK(sfci_dot(ctx, v_new, w_new, &beta_new2));
    double beta_new = std::sqrt(beta_new2);
    const double beta_one = beta_new;
    CK(sfci_scale(ctx, v_new, 1.0 / beta_new));
    CK(sfci_scale(ctx, w_new, 1.0 / beta_new));
    double c = 1.0, c_old = 1.0, s = 0.0, s_old = 0.0, eta = 1.0;
    int it = 0;
    while (it < iters) {
        const double beta = beta_new;
        CK(sfci_vec_copy(v_old, v));
        CK(sfci_vec_copy(v, v_new));
        CK(sfci_vec_copy(w, w_new));
        CK(sfci_hv(ctx, w, v_new, nd, lambda, 1.0, &diag));
        CK(sfci_add_two(ctx, v_new, v_old, -beta));
        double alpha = 0.0;
        CK(sfci_dot(ctx, v_new, w, &alpha));
        CK(sfci_add_two(ctx, v_new, v, -alpha));
        CK(sfci_vec_copy(w_new, v_new));
        CK(sfci_dot(ctx, v_new, w_new, &beta_new2));
        beta_new = std::sqrt(beta_new2);
        CK(sfci_scale(ctx, v_new, 1.0 / beta_new));
        CK(sfci_scale(ctx, w_new, 1.0 / beta_new));
        const double r2 = s * alpha + c * c_old * beta;
        const double r3 = s_old * beta;
        const double r1_hat = c * alpha - c_old * s * beta;
        const double r1 = std::sqrt(r1_hat * r1_hat + beta_new * beta_new);
        c_old = c;
        s_old = s;
        c = r1_hat / r1;
        s = beta_new / r1;
        CK(sfci_vec_copy(p_oold, p_old));
        CK(sfci_vec_copy(p_old, pv));
        CK(sfci_add_three(ctx, pv, w, p_old, p_oold, 1.0 / r1, -r2 / r1, -r3 / r1));
        CK(sfci_add_two(ctx, x, pv, beta_one * c * eta));
        residualNorm2 *= s * s;
        if (residualNorm2 < tol) break;
        eta = -s * eta;
        ++it;
    }
    CK(sfci_vec_copy(result, x));
    if (iters_done) *iters_done = it;
    if (res2_out) *res2_out = residualNorm2;
    return sfci_sync(ctx);
}
