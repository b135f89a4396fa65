// Ground-state driver: sparse_krylov_solve_eigenval (src/sparse_solver.cpp:246-548)
// restated in host C++ on top of the library's C-ABI.  Every vector stays on the GPU;
// the host only holds the Nl x Nl projected matrices.
#include <chrono>
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

#define CK(x)                      \
    do {                           \
        int rc_ = (x);             \
        if (rc_ != SFCI_OK) return rc_; \
    } while (0)

using clk = std::chrono::steady_clock;
double ms_since(clk::time_point t) { return std::chrono::duration<double, std::milli>(clk::now() - t).count(); }

}  // namespace

extern "C" int sfci_ground_state(sfci_ctx* ctx, uint64_t init_det, const sfci_lanczos_params_t* p, double* ritz,
                                 double* e_trunc, sfci_lanczos_result_t* res, sfci_vec* v_out) {
    if (!ctx || !p || !ritz || !e_trunc || !res || p->nd < 1 || p->nl < 1 || p->num_restart < 1 || p->hpsi_cap < 1)
        return SFCI_E_ARG;
    const int nl = p->nl;
    const int64_t nd = p->nd, cap = p->hpsi_cap;
    auto t_start = clk::now();
    *res = sfci_lanczos_result_t{};
    VecGuard g;
    auto mk = [&](int64_t c, sfci_vec** out) {
        int rc = sfci_vec_create(ctx, c, out);
        if (rc == SFCI_OK) g.v.push_back(*out);
        return rc;
    };
    std::vector<sfci_vec*> K(nl);
    for (int i = 0; i < nl; ++i) CK(mk(nd, &K[i]));
    sfci_vec *v, *u, *temp, *temp_u, *v_next;
    CK(mk(nd, &v));
    CK(mk(cap, &u));
    CK(mk(cap, &temp));
    CK(mk(cap, &temp_u));
    CK(mk(nd, &v_next));
    double one = 1.0;
    CK(sfci_vec_upload(v, &init_det, &one, 1));
    // screening: the unscreened energy readout judges every restart (no reference counterpart)
    int smode = SFCI_SCREEN_OFF;
    sfci_mlp* snet = nullptr;
    double skeep = 0.0, seps = 0.0;
    CK(sfci_get_screen(ctx, &smode, &snet, &skeep, &seps));
    bool screening = smode != SFCI_SCREEN_OFF;
    struct ScreenRestore {
        sfci_ctx* ctx;
        int mode;
        sfci_mlp* net;
        double keep, eps;
        ~ScreenRestore() {
            if (mode != SFCI_SCREEN_OFF) sfci_set_screen(ctx, mode, net, keep, eps);
        }
    } restore{ctx, smode, snet, skeep, seps};
    sfci_vec* v_best = nullptr;
    if (screening) CK(mk(nd, &v_best));
    double best_e = INFINITY;
    bool verifying = false;
    const bool judged = screening;

    const double deltaE = p->delta_e;
    double previousE = p->e_prev_init, nearestE = -1000000.0;
    std::vector<double> KM(nl * nl, 0.0), S(nl * nl, 0.0);
    auto hv = [&](const sfci_vec* in, sfci_vec* out, double* alpha) {
        auto t = clk::now();
        int rc = sfci_hv(ctx, in, out, nd, 0.0, 1.0, alpha);
        res->ms_hv += ms_since(t);
        res->n_hv += 1;
        return rc;
    };
    int r = 0;
    for (r = 0; r < p->num_restart; ++r) {
        if (p->verbose) {
            printf("\n*****************block %d starts*****************\n", r);
            fflush(stdout);
        }
        CK(sfci_vec_copy(K[0], v));
        double alpha = 0.0, beta = 0.0;
        CK(hv(v, u, &alpha));
        if (p->verbose) {
            printf("H * phi_0 len:%lld\n", (long long)sfci_vec_len(u));
            fflush(stdout);
        }
        KM[0] = alpha;
        for (int i = 1; i < nl; ++i) {
            auto t_it = clk::now();
            CK(sfci_vec_copy(temp, u));
            CK(sfci_add_two(ctx, temp, v, -alpha));
            CK(sfci_shrink_from(ctx, v_next, temp));
            double gamma = 0.0;
            CK(sfci_norm(ctx, v_next, &gamma));
            beta = gamma;
            CK(sfci_scale(ctx, v_next, 1.0 / gamma));
            CK(hv(v_next, temp_u, &alpha));
            KM[i * nl + i] = alpha;
            for (int j = 0; j < i; ++j) {
                double d = 0.0;
                CK(sfci_dot(ctx, K[j], temp_u, &d));
                KM[j * nl + i] = KM[i * nl + j] = d;
            }
            CK(sfci_add_two(ctx, temp_u, v, -beta));
            CK(sfci_vec_copy(u, temp_u));
            CK(sfci_vec_copy(v, v_next));
            CK(sfci_vec_copy(K[i], v));
            if (p->verbose) {
                sfci_hv_stats_t hs{};
                sfci_hv_stats(ctx, &hs);
                printf("iter %d: len %lld alpha %.12f beta %.12f time %.3f ms | hv %.1f ms (gen %.1f sort %.1f "
                       "reduce %.1f) passes %llu unique %llu out %llu fallback %llu hb %d split %llu\n",
                       i, (long long)sfci_vec_len(v), alpha, beta, ms_since(t_it), hs.ms_total, hs.ms_generate,
                       hs.ms_sort, hs.ms_reduce, (unsigned long long)hs.n_passes, (unsigned long long)hs.n_unique,
                       (unsigned long long)hs.n_out, (unsigned long long)hs.n_fallback, hs.hash_bits,
                       (unsigned long long)hs.n_split);
                fflush(stdout);
            }
        }
        for (int i = 0; i < nl; ++i)
            for (int j = i; j < nl; ++j) {
                double d = 0.0;
                CK(sfci_dot(ctx, K[i], K[j], &d));
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
        CK(sfci_lincomb(ctx, K.data(), c.data(), nl, v));
        double nrm = 0.0;
        CK(sfci_norm(ctx, v, &nrm));
        CK(sfci_scale(ctx, v, 1.0 / nrm));
        double et = 0.0;
        CK(sfci_energy(ctx, v, &et));
        // Unscreened restart: the reference's test on the Ritz value.  With screening, every
        // restart is judged by the unscreened readout et and the lowest-energy vector is kept:
        // a screened restart whose energy rises is rejected (best vector back, unscreened from
        // here on); one that settles (drop < deltaE) hands over to an unscreened restart, which
        // stops the solve unless it lowers the energy by deltaE or more.  The Ritz values of
        // screened restarts are not compared with unscreened ones.
        bool stop = previousE - ev[nearest] < deltaE;
        bool reset_prev = false;
        if (judged) {
            const double prev_best = best_e;
            if (et < best_e) {
                best_e = et;
                CK(sfci_vec_copy(v_best, v));
            }
            if (screening) {
                res->n_screened += 1;
                stop = false;
                if (et > prev_best + 1e-9) {
                    CK(sfci_vec_copy(v, v_best));
                    et = best_e;
                    res->n_reverts += 1;
                    if (p->verbose) printf("screened restart rejected: energy rose; continuing unscreened\n");
                } else if (prev_best - et < deltaE) {
                    verifying = true;
                    if (p->verbose) printf("screened solve settled; confirming with an unscreened restart\n");
                }
                if (res->n_reverts > 0 || verifying) {
                    CK(sfci_set_screen(ctx, SFCI_SCREEN_OFF, nullptr, 0.0, 0.0));
                    screening = false;
                    reset_prev = true;   // the first unscreened restart compares with +inf
                    nearestE = -1000000.0;   // and takes the lowest Ritz value, not the one nearest a
                                             // screened restart's
                }
            } else if (verifying) {
                verifying = false;
                stop = prev_best - et < deltaE;
                if (!stop) {
                    res->n_reverts += 1;
                    if (p->verbose) printf("unscreened restart lowered the energy further; continuing unscreened\n");
                }
            }
        }
        ritz[r] = ev[nearest];
        e_trunc[r] = et;
        if (p->ms_restart) p->ms_restart[r] = ms_since(t_start);
        if (p->verbose) {
            printf("energy Psi_res after truncation:%.12f\n", et);
            printf("gsaes energy:%.12f\n", ev[nearest]);
            printf("*****************block %d ends*****************\n", r);
            fflush(stdout);
        }
        std::fill(KM.begin(), KM.end(), 0.0);
        for (int i = 0; i < nl; ++i) CK(sfci_vec_clear(K[i]));
        if (stop) {
            ++r;
            break;
        }
        previousE = reset_prev ? INFINITY : ev[nearest];
    }
    if (judged) CK(sfci_vec_copy(v, v_best));   // the lowest-energy vector of the solve
    double ef = 0.0;
    CK(sfci_energy(ctx, v, &ef));
    if (p->verbose) printf("final e:%.12f\n", ef);
    if (v_out) CK(sfci_vec_copy(v_out, v));
    CK(sfci_sync(ctx));
    res->n_restarts = r;
    res->e_final = ef;
    res->ms_total = ms_since(t_start);
    return SFCI_OK;
}
