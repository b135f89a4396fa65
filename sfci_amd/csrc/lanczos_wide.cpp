// Ground-state driver on 128-bit keys: sparse_krylov_solve_eigenval
// (src/sparse_solver.cpp:246-548) restated, as sfci_ground_state does for 64-bit keys, on
// top of the sfci_wide_* C-ABI.  The steps, constants and stopping test are those of the
// 64-bit driver without candidate screening; every vector stays on the GPU and the host
// holds only the Nl x Nl projected matrices.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/sfci_gpu.h"
#include "gsaes.h"

namespace {

struct WVecGuard {
    std::vector<sfci_wvec*> v;
    ~WVecGuard() {
        for (auto* p : v) sfci_wide_vec_destroy(p);
    }
};

#define CKW(x)                          \
    do {                                \
        int rc_ = (x);                  \
        if (rc_ != SFCI_OK) return rc_; \
    } while (0)

using wclk = std::chrono::steady_clock;
double ms_since_w(wclk::time_point t) { return std::chrono::duration<double, std::milli>(wclk::now() - t).count(); }

}  // namespace

extern "C" int sfci_wide_ground_state(sfci_wctx* ctx, const uint64_t* init_det, const sfci_lanczos_params_t* p,
                                      double* ritz, double* e_trunc, sfci_lanczos_result_t* res, sfci_wvec* v_out) {
    if (!ctx || !init_det || !p || !ritz || !e_trunc || !res || p->nd < 1 || p->nl < 1 || p->num_restart < 1 ||
        p->hpsi_cap < 1)
        return SFCI_E_ARG;
    const int nl = p->nl;
    const int64_t nd = p->nd, cap = p->hpsi_cap;
    auto t_start = wclk::now();
    *res = sfci_lanczos_result_t{};
    WVecGuard g;
    auto mk = [&](int64_t c, sfci_wvec** out) {
        int rc = sfci_wide_vec_create(ctx, c, out);
        if (rc == SFCI_OK) g.v.push_back(*out);
        return rc;
    };
    std::vector<sfci_wvec*> K(nl);
    for (int i = 0; i < nl; ++i) CKW(mk(nd, &K[i]));
    sfci_wvec *v, *u, *temp, *temp_u, *v_next;
    CKW(mk(nd, &v));
    CKW(mk(cap, &u));
    CKW(mk(cap, &temp));
    CKW(mk(cap, &temp_u));
    CKW(mk(nd, &v_next));
    double one = 1.0;
    CKW(sfci_wide_vec_upload(v, init_det, &one, 1));
    const double deltaE = p->delta_e;
    double previousE = p->e_prev_init, nearestE = -1000000.0;
    std::vector<double> KM(nl * nl, 0.0), S(nl * nl, 0.0);
    auto hv = [&](const sfci_wvec* in, sfci_wvec* out, double* alpha) {
        auto t = wclk::now();
        int rc = sfci_wide_hv(ctx, in, out, nd, 0.0, 1.0, alpha);
        res->ms_hv += ms_since_w(t);
        res->n_hv += 1;
        return rc;
    };
    int r = 0;
    for (r = 0; r < p->num_restart; ++r) {
        if (p->verbose) {
            printf("\n*****************block %d starts*****************\n", r);
            fflush(stdout);
        }
        CKW(sfci_wide_vec_copy(K[0], v));
        double alpha = 0.0, beta = 0.0;
        CKW(hv(v, u, &alpha));
        KM[0] = alpha;
        for (int i = 1; i < nl; ++i) {
            auto t_it = wclk::now();
            CKW(sfci_wide_vec_copy(temp, u));
            CKW(sfci_wide_add_two(ctx, temp, v, -alpha));
            CKW(sfci_wide_shrink_from(ctx, v_next, temp));
            double gamma = 0.0;
            CKW(sfci_wide_norm(ctx, v_next, &gamma));
            beta = gamma;
            CKW(sfci_wide_scale(ctx, v_next, 1.0 / gamma));
            CKW(hv(v_next, temp_u, &alpha));
            KM[i * nl + i] = alpha;
            for (int j = 0; j < i; ++j) {
                double d = 0.0;
                CKW(sfci_wide_dot(ctx, K[j], temp_u, &d));
                KM[j * nl + i] = KM[i * nl + j] = d;
            }
            CKW(sfci_wide_add_two(ctx, temp_u, v, -beta));
            CKW(sfci_wide_vec_copy(u, temp_u));
            CKW(sfci_wide_vec_copy(v, v_next));
            CKW(sfci_wide_vec_copy(K[i], v));
            if (p->verbose) {
                sfci_hv_stats_t hs{};
                sfci_wide_hv_stats(ctx, &hs);
                printf("iter %d: len %lld alpha %.12f beta %.12f time %.3f ms | hv %.1f ms (gen %.1f sort %.1f "
                       "reduce %.1f) passes %llu unique %llu out %llu\n",
                       i, (long long)sfci_wide_vec_len(v), alpha, beta, ms_since_w(t_it), hs.ms_total,
                       hs.ms_generate, hs.ms_sort, hs.ms_reduce, (unsigned long long)hs.n_passes,
                       (unsigned long long)hs.n_unique, (unsigned long long)hs.n_out);
                fflush(stdout);
            }
        }
        for (int i = 0; i < nl; ++i)
            for (int j = i; j < nl; ++j) {
                double d = 0.0;
                CKW(sfci_wide_dot(ctx, K[i], K[j], &d));
                S[i * nl + j] = S[j * nl + i] = d;
            }
        std::vector<double> ev, evec;
        if (sfci_dense::gsaes(nl, KM, S, ev, evec)) return SFCI_E_ARG;
        int nearest = 0;
        double maxGap = 0.0;
        for (int i = 0; i < nl; ++i) {
            double gap = std::fabs(nearestE - ev[i]);
            if (i == 0 || gap < maxGap) {
                maxGap = gap;
                nearest = i;
            }
        }
        nearestE = ev[nearest];
        std::vector<double> c(nl);
        for (int i = 0; i < nl; ++i) c[i] = evec[i * nl + nearest];
        CKW(sfci_wide_lincomb(ctx, K.data(), c.data(), nl, v));
        double nrm = 0.0;
        CKW(sfci_wide_norm(ctx, v, &nrm));
        CKW(sfci_wide_scale(ctx, v, 1.0 / nrm));
        double et = 0.0;
        CKW(sfci_wide_energy(ctx, v, &et));
        const bool stop = previousE - ev[nearest] < deltaE;
        ritz[r] = ev[nearest];
        e_trunc[r] = et;
        if (p->ms_restart) p->ms_restart[r] = ms_since_w(t_start);
        if (p->verbose) {
            printf("energy Psi_res after truncation:%.12f\n", et);
            printf("gsaes energy:%.12f\n", ev[nearest]);
            printf("*****************block %d ends*****************\n", r);
            fflush(stdout);
        }
        std::fill(KM.begin(), KM.end(), 0.0);
        for (int i = 0; i < nl; ++i) CKW(sfci_wide_vec_clear(K[i]));
        if (stop) {
            ++r;
            break;
        }
        previousE = ev[nearest];
    }
    double ef = 0.0;
    CKW(sfci_wide_energy(ctx, v, &ef));
    if (p->verbose) printf("final e:%.12f\n", ef);
    if (v_out) CKW(sfci_wide_vec_copy(v_out, v));
    CKW(sfci_wide_sync(ctx));
    res->n_restarts = r;
    res->e_final = ef;
    res->ms_total = ms_since_w(t_start);
    return SFCI_OK;
}
