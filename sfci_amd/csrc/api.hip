// C-ABI of libsfci_hip.so: context, integral loading, device vectors, operations.
#include <cmath>
#include <cstdio>
#include <exception>

#include "../../include/sfci_gpu.h"
#include "internal.h"

using namespace sfci;

namespace {

uint64_t binom_u64(int n, int k) {
    if (k < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    unsigned __int128 r = 1;
    for (int i = 0; i < k; ++i) r = r * (unsigned)(n - i) / (unsigned)(i + 1);
    return (uint64_t)r;
}

struct Entry {
    double c;
    int m, n, p, q;
};

// ReadElements (src/functions.cpp:297-395): up to 7 blocks of "count" + count lines of
// "coeff m n p q"; stops early after a block of exactly one line.
int read_in_file(const char* path, std::vector<Entry>& out) {
    FILE* fp = fopen(path, "r");
    if (!fp) return SFCI_E_IO;
    for (int blk = 0; blk < 7; ++blk) {
        int num = 0;
        if (fscanf(fp, "%d", &num) != 1) break;
        for (int i = 0; i < num; ++i) {
            Entry e;
            if (fscanf(fp, "%lf %d %d %d %d", &e.c, &e.m, &e.n, &e.p, &e.q) != 5) {
                fclose(fp);
                return SFCI_E_ARG;
            }
            out.push_back(e);
        }
        if (num == 1) break;
    }
    fclose(fp);
    return SFCI_OK;
}

// max_so: 64 for the 64-bit-key entry points, 128 for the 128-bit-key (sfci_wide_*) ones.
int build_ctx(int n_so, int n_elec, const std::vector<Entry>& ent, int device, sfci_ctx** out, std::string& err,
              int max_so = 64) {
    if (n_so < 2 || n_so > max_so || (n_so & 1) || n_elec < 2 || (n_elec & 1) || n_elec / 2 > n_so / 2 ||
        n_elec / 2 > SFCI_MAX_K) {
        err = max_so == 64 ? "unsupported n_so / n_elec (n_so even <= 64, n_elec even)"
                           : "unsupported n_so / n_elec (n_so even <= 128, n_elec even, n_elec <= 64)";
        return SFCI_E_ARG;
    }
    const int N = n_so / 2;
    const int k = n_elec / 2;
    const size_t N2 = (size_t)N * N, N4 = N2 * N2;
    std::vector<double> h(2 * N2, 0.0), g(4 * N4, 0.0);
    double constant = 0.0;
    for (const Entry& e : ent) {
        if (e.p == 0 && e.q == 0) {
            if (e.m == 0 && e.n == 0) {
                constant = e.c;
                continue;
            }
            int a = e.m - 2, b = e.n - 2;
            if (a < 0 || b < 0 || a >= n_so || b >= n_so) {
                err = "one-body index out of range";
                return SFCI_E_ARG;
            }
            if ((a ^ b) & 1) {
                err = "spin-forbidden one-body entry";
                return SFCI_E_ARG;
            }
            h[(size_t)(a & 1) * N2 + (size_t)(a >> 1) * N + (b >> 1)] += e.c;
        } else {
            int a = e.m - 2, b = e.n - 2, p = e.p - 2, q = e.q - 2;
            if (a < 0 || b < 0 || p < 0 || q < 0 || a >= n_so || b >= n_so || p >= n_so || q >= n_so) {
                err = "two-body index out of range";
                return SFCI_E_ARG;
            }
            if (((a ^ b) & 1) || ((p ^ q) & 1)) {
                err = "spin-forbidden two-body entry";
                return SFCI_E_ARG;
            }
            size_t blk = (size_t)(((a & 1) << 1) | (p & 1));
            g[blk * N4 + (((size_t)(a >> 1) * N + (b >> 1)) * N + (p >> 1)) * N + (q >> 1)] += e.c;
        }
    }
    sfci_ctx* X = new sfci_ctx();
    Ctx& c = X->c;
    try {
        HIPCHK(hipSetDevice(device));
        c.device = device;
        HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        noalloc_setup(device);
        c.n_so = n_so;
        c.n_orb = N;
        c.n_elec = n_elec;
        c.h_binom.assign(65 * 33, 0);
        for (int n = 0; n <= 64; ++n)
            for (int t = 0; t <= 32; ++t) c.h_binom[n * 33 + t] = binom_u64(n, t);
        const int v = N - k;
        std::vector<uint16_t> op, vp;
        for (int i = 0; i < k; ++i)
            for (int j = i + 1; j < k; ++j) op.push_back((uint16_t)(i | (j << 8)));
        for (int i = 0; i < v; ++i)
            for (int j = i + 1; j < v; ++j) vp.push_back((uint16_t)(i | (j << 8)));
        std::vector<uint16_t> sd;
        for (int x = 0; x < k * v; ++x) sd.push_back((uint16_t)((x / v) | ((x % v) << 8)));
        if (op.empty()) op.push_back(0);
        if (vp.empty()) vp.push_back(0);
        if (sd.empty()) sd.push_back(0);
        // One block for all four spin cases when the file is spin-free (RHF): the opposite-
        // spin blocks agree entry for entry, and the same-spin blocks agree with them on
        // every entry the Slater-Condon rules read (same-spin (a i|b j) is only read with
        // a != b and i != j; the converter leaves a == b or i == j out of those blocks).
        bool one_block = true;
        for (size_t t = 0; t < N4 && one_block; ++t) one_block = g[N4 + t] == g[2 * N4 + t];
        for (size_t a = 0; a < (size_t)N && one_block; ++a)
            for (size_t i = 0; i < (size_t)N && one_block; ++i)
                for (size_t b = 0; b < (size_t)N && one_block; ++b)
                    for (size_t j = 0; j < (size_t)N; ++j) {
                        if (a == b || i == j) continue;
                        size_t t = ((a * N + i) * N + b) * N + j;
                        if (g[t] != g[N4 + t] || g[3 * N4 + t] != g[N4 + t]) {
                            one_block = false;
                            break;
                        }
                    }
        // opposite-spin blocks with the spin-1 virtual index fastest: the layout the
        // opposite-spin doubles read with consecutive lanes on consecutive spin-1 virtuals
        const size_t n_gt = one_block ? 1 : 2;
        std::vector<double> gt(n_gt * N4);
        for (size_t a = 0; a < (size_t)N; ++a)
            for (size_t i = 0; i < (size_t)N; ++i)
                for (size_t j = 0; j < (size_t)N; ++j)
                    for (size_t b = 0; b < (size_t)N; ++b) {
                        const size_t t = ((a * N + i) * N + j) * N + b;
                        gt[t] = g[N4 + ((a * N + i) * N + b) * N + j];                           // (a i|b j), (0,1)
                        if (n_gt == 2) gt[N4 + t] = g[2 * N4 + ((b * N + j) * N + a) * N + i];   // (b j|a i), (1,0)
                    }
        if (one_block) {
            std::vector<double> g1(g.begin() + N4, g.begin() + 2 * N4);
            g.swap(g1);
        }
        c.g_one_block = one_block;
        HIPCHK(hipMalloc(&c.d_h, h.size() * 8));
        HIPCHK(hipMalloc(&c.d_g, g.size() * 8));
        HIPCHK(hipMalloc(&c.d_gt, gt.size() * 8));
        HIPCHK(hipMemcpy(c.d_gt, gt.data(), gt.size() * 8, hipMemcpyHostToDevice));
        // the device's table: rows n <= 32 in 32 bits (C(32, 16) < 2^32)
        std::vector<uint32_t> b32(33 * 33);
        for (int i = 0; i < 33 * 33; ++i) b32[i] = (uint32_t)c.h_binom[i];
        HIPCHK(hipMalloc(&c.d_binom, b32.size() * 4));
        HIPCHK(hipMalloc(&c.d_occ_pairs, op.size() * 2));
        HIPCHK(hipMalloc(&c.d_vir_pairs, vp.size() * 2));
        HIPCHK(hipMalloc(&c.d_sdec, sd.size() * 2));
        HIPCHK(hipMemcpy(c.d_sdec, sd.data(), sd.size() * 2, hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&c.d_acc, 16 * 8));
        HIPCHK(hipMalloc(&c.d_flags, 16));
        HIPCHK(hipMalloc(&c.d_counter, 8 * 8));
        HIPCHK(hipMemcpy(c.d_h, h.data(), h.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c.d_g, g.data(), g.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c.d_binom, b32.data(), b32.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c.d_occ_pairs, op.data(), op.size() * 2, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c.d_vir_pairs, vp.data(), vp.size() * 2, hipMemcpyHostToDevice));
        HIPCHK(hipMemset(c.d_acc, 0, 16 * 8));
        HIPCHK(hipMemset(c.d_flags, 0, 16));
        HIPCHK(hipMemset(c.d_counter, 0, 8 * 8));
        HamDev& H = c.ham;
        H.n_orb = N;
        H.k = k;
        H.v = v;
        H.dim = binom_u64(N, k);
        H.h = c.d_h;
        H.g = c.d_g;
        H.g_stride = one_block ? 0 : N4;
        H.gt = c.d_gt;
        H.gt_stride = n_gt == 2 ? N4 : 0;
        H.binom = c.d_binom;
        H.occ_pairs = c.d_occ_pairs;
        H.vir_pairs = c.d_vir_pairs;
        H.constant = constant;
        H.n_sa = (uint64_t)k * v;
        H.n_opair = (uint32_t)(k * (k - 1) / 2);
        H.n_vpair = (uint32_t)(v * (v - 1) / 2);
        H.n_daa = (uint64_t)H.n_opair * H.n_vpair;
        H.n_dab = H.n_sa * H.n_sa;
        H.n_conn = 1 + 2 * H.n_sa + 2 * H.n_daa + H.n_dab;
        if (H.n_vpair == 0) H.n_vpair = 1;   // divisor guard; n_daa is 0 then
        size_t free_b = 0, total_b = 0;
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        c.ws_limit = (size_t)(free_b * 0.55);
    } catch (const HipError& e) {
        err = e.what();
        delete X;
        return e.code;
    }
    *out = X;
    return SFCI_OK;
}

}  // namespace

extern "C" {

int sfci_create(const char* in_path, int n_so, int n_elec, int device, sfci_ctx** out) {
    if (!in_path || !out) return SFCI_E_ARG;
    *out = nullptr;
    std::vector<Entry> ent;
    int rc = read_in_file(in_path, ent);
    if (rc) return rc;
    std::string err;
    rc = build_ctx(n_so, n_elec, ent, device, out, err);
    if (rc) fprintf(stderr, "sfci_create: %s\n", err.c_str());
    return rc;
}

int sfci_create_from_entries(int n_so, int n_elec, int64_t n_entries, const double* coeff, const int32_t* idx,
                             int device, sfci_ctx** out) {
    if (!out || n_entries < 0 || (n_entries && (!coeff || !idx))) return SFCI_E_ARG;
    *out = nullptr;
    std::vector<Entry> ent(n_entries);
    for (int64_t i = 0; i < n_entries; ++i)
        ent[i] = Entry{coeff[i], idx[4 * i], idx[4 * i + 1], idx[4 * i + 2], idx[4 * i + 3]};
    std::string err;
    int rc = build_ctx(n_so, n_elec, ent, device, out, err);
    if (rc) fprintf(stderr, "sfci_create_from_entries: %s\n", err.c_str());
    return rc;
}

void sfci_destroy(sfci_ctx* X) {
    if (!X) return;
    Ctx& c = X->c;
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    DevBuf* bufs[] = {&c.cand_k[0], &c.cand_k[1], &c.cand_v[0], &c.cand_v[1], &c.src_cnt, &c.src_off,
                      &c.src_k, &c.src_v, &c.dterm, &c.bucket_bounds, &c.lin_k[0], &c.lin_k[1], &c.lin_v[0],
                      &c.lin_v[1], &c.part_meta[0], &c.part_meta[1], &c.part_start[0], &c.part_start[1],
                      &c.part_len[0], &c.part_len[1], &c.part_hbase[0], &c.part_hbase[1], &c.part_hstride[0],
                      &c.part_hstride[1], &c.part_region[0], &c.part_region[1], &c.part_base, &c.part_dig, &c.stage_k, &c.stage_v, &c.tmp_k[0], &c.tmp_k[1],
                      &c.tmp_v[0], &c.tmp_v[1], &c.sel_a, &c.sel_b, &c.sel_c, &c.scratch_hist,
                      &c.scratch_off, &c.scratch_small, &c.rowhist, &c.plan_k[0], &c.plan_k[1], &c.plan_v[0], &c.plan_v[1],
                      &c.plan_fl, &c.plan_pos, &c.plan_first, &c.plan_str, &c.plan_cnt};
    for (DevBuf* b : bufs) b->release();
    for (auto& b : c.scratch_scan) b.release();
    void* ptrs[] = {c.d_h, c.d_g, c.d_gt, c.d_binom, c.d_occ_pairs, c.d_vir_pairs, c.d_sdec, c.d_acc, c.d_flags, c.d_counter};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete X;
}

const char* sfci_last_error(const sfci_ctx* X) { return X ? X->c.last_error.c_str() : "null context"; }

int sfci_info(const sfci_ctx* X, sfci_info_t* info) {
    if (!X || !info) return SFCI_E_ARG;
    const Ctx& c = X->c;
    info->n_so = c.n_so;
    info->n_elec = c.n_elec;
    info->string_dim = c.ham.dim;
    info->n_conn = c.ham.n_conn;
    info->e_core = c.ham.constant;
    return SFCI_OK;
}

int sfci_set_cut(sfci_ctx* X, double cut) {
    if (!X || !(cut >= 0.0)) return SFCI_E_ARG;
    X->c.cut = cut;
    return SFCI_OK;
}

int sfci_set_workspace(sfci_ctx* X, uint64_t bytes) {
    if (!X || bytes < (1u << 20)) return SFCI_E_ARG;
    X->c.ws_limit = bytes;
    return SFCI_OK;
}

int sfci_hv_stats(const sfci_ctx* X, sfci_hv_stats_t* s) {
    if (!X || !s) return SFCI_E_ARG;
    const auto& t = X->c.st;
    s->n_src = t.n_src;
    s->n_candidates = t.n_candidates;
    s->n_unique = t.n_unique;
    s->n_out = t.n_out;
    s->n_passes = t.n_passes;
    s->ms_generate = t.ms_generate;
    s->ms_sort = t.ms_sort;
    s->ms_reduce = t.ms_reduce;
    s->ms_select = t.ms_select;
    s->ms_total = t.ms_total;
    s->n_fallback = t.n_fallback;
    s->hash_bits = t.hash_bits;
    s->n_kept = t.n_kept;
    s->screen_eps = t.screen_eps;
    s->ms_score = t.ms_score;
    s->n_split = t.n_split;
    s->n_str0 = t.n_str0;
    s->n_str1 = t.n_str1;
    return SFCI_OK;
}

int sfci_set_precision(sfci_ctx* X, int precision) {
    if (!X || (precision != SFCI_PREC_FP64 && precision != SFCI_PREC_MIXED)) return SFCI_E_ARG;
    X->c.precision = precision;
    return SFCI_OK;
}

int sfci_set_merge_mode(sfci_ctx* X, int mode, int hash_bits) {
    if (!X || mode < 0 || mode > 4 || (mode >= 2 && (hash_bits < 1 || hash_bits > 16))) return SFCI_E_ARG;
    X->c.merge_mode = mode;
    X->c.merge_hash_bits = hash_bits;
    return SFCI_OK;
}

int sfci_set_gen_tables(sfci_ctx* X, int mode) {
    if (!X || mode < 0 || mode > 2) return SFCI_E_ARG;
    X->c.gen_tables = mode;
    return SFCI_OK;
}

int sfci_kernel_stats(sfci_ctx* X, sfci_kernel_stat_t* out, int n, int reset) {
    if (!X || n < 0 || (n && !out)) return SFCI_E_ARG;
    Ctx* c = &X->c;
    return guarded(c, [&] {
        kt_resolve(c);
        for (int i = 0; i < n && i < K_NUM; ++i) {
            out[i].ms = c->kclock.stat[i].ms;
            out[i].launches = c->kclock.stat[i].launches;
            out[i].bytes = c->kclock.stat[i].bytes;
        }
        if (reset)
            for (auto& s : c->kclock.stat) s = KernelStat{};
    });
}

const char* sfci_kernel_name(int id) {
    static const char* names[K_NUM] = {"k_generate",      "k_rowgroup_hist+k_count_passes",
                                       "k_part_hist<DigitHash>", "k_part_scatter<DigitHash>",
                                       "k_radix_hist<DigitPlain>", "k_radix_scatter<DigitPlain>",
                                       "scan",            "k_bucket_bounds",
                                       "k_bucket_reduce", "k_reduce",
                                       "select",          "vecops",
                                       "k_energy",        "k_part_scatter_cur<DigitHash>",
                                       "k_mlp_fwd",       "screen_select",
                                       "mlp_train_step",  "k_ss_table"};
    return (id >= 0 && id < K_NUM) ? names[id] : "";
}

int sfci_sync(sfci_ctx* X) {
    if (!X) return SFCI_E_ARG;
    return guarded(&X->c, [&] { HIPCHK(hipStreamSynchronize(X->c.stream)); });
}

int sfci_vec_create(sfci_ctx* X, int64_t cap, sfci_vec** out) {
    if (!X || !out || cap < 0) return SFCI_E_ARG;
    *out = nullptr;
    sfci_vec* V = new sfci_vec();
    V->v.ctx = &X->c;
    V->v.cap = cap;
    int rc = guarded(&X->c, [&] {
        size_t n = cap > 0 ? (size_t)cap : 1;
        HIPCHK(hipMalloc(&V->v.k, n * 8));
        HIPCHK(hipMalloc(&V->v.v, n * 8));
    });
    if (rc) {
        if (V->v.k) (void)hipFree(V->v.k);
        if (V->v.v) (void)hipFree(V->v.v);
        delete V;
        return rc;
    }
    *out = V;
    return SFCI_OK;
}

void sfci_vec_destroy(sfci_vec* V) {
    if (!V) return;
    (void)hipSetDevice(V->v.ctx->device);
    (void)hipStreamSynchronize(V->v.ctx->stream);
    if (V->v.k) (void)hipFree(V->v.k);
    if (V->v.v) (void)hipFree(V->v.v);
    delete V;
}

int64_t sfci_vec_len(const sfci_vec* V) { return V ? V->v.len : -1; }
int64_t sfci_vec_cap(const sfci_vec* V) { return V ? V->v.cap : -1; }

int sfci_vec_upload(sfci_vec* V, const uint64_t* keys, const double* vals, int64_t n) {
    if (!V || n < 0 || n > V->v.cap || (n && (!keys || !vals))) return SFCI_E_ARG;
    Ctx* c = V->v.ctx;
    for (int64_t i = 0; i < n; ++i) {
        if (i > 0 && keys[i] <= keys[i - 1]) {
            c->last_error = "sfci_vec_upload: keys not strictly ascending";
            return SFCI_E_ARG;
        }
        uint64_t d = keys[i];
        if (popc64(d & 0x5555555555555555ull) != c->ham.k || popc64(d & 0xaaaaaaaaaaaaaaaaull) != c->ham.k ||
            (c->n_so < 64 && (d >> c->n_so))) {
            c->last_error = "sfci_vec_upload: key " + std::to_string(d) + " does not match n_so / n_elec";
            return SFCI_E_ARG;
        }
    }
    return guarded(c, [&] {
        if (n) {
            HIPCHK(hipMemcpyAsync(V->v.k, keys, n * 8, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(V->v.v, vals, n * 8, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        V->v.len = n;
    });
}

int sfci_vec_download(const sfci_vec* V, uint64_t* keys, double* vals, int64_t max_n, int64_t* n_out) {
    if (!V || max_n < 0) return SFCI_E_ARG;
    Ctx* c = V->v.ctx;
    int64_t n = V->v.len < max_n ? V->v.len : max_n;
    if (n && (!keys || !vals)) return SFCI_E_ARG;
    return guarded(c, [&] {
        if (n) {
            HIPCHK(hipMemcpyAsync(keys, V->v.k, n * 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(vals, V->v.v, n * 8, hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        if (n_out) *n_out = n;
    });
}

int sfci_vec_device_ptrs(sfci_vec* V, uint64_t** keys, double** vals) {
    if (!V || !keys || !vals) return SFCI_E_ARG;
    *keys = V->v.k;
    *vals = V->v.v;
    return SFCI_OK;
}

int sfci_vec_set_len(sfci_vec* V, int64_t len) {
    if (!V || len < 0 || len > V->v.cap) return SFCI_E_ARG;
    V->v.len = len;
    return SFCI_OK;
}

int sfci_vec_copy(sfci_vec* D, const sfci_vec* S) {
    if (!D || !S || D->v.ctx != S->v.ctx || S->v.len > D->v.cap) return SFCI_E_ARG;
    Ctx* c = D->v.ctx;
    return guarded(c, [&] {
        dev_copy(c, D->v.k, D->v.v, S->v.k, S->v.v, (size_t)S->v.len);
        D->v.len = S->v.len;
    });
}

int sfci_vec_clear(sfci_vec* V) {
    if (!V) return SFCI_E_ARG;
    V->v.len = 0;
    return SFCI_OK;
}

int sfci_hv(sfci_ctx* X, const sfci_vec* in, sfci_vec* out, int64_t nd, double lambda, double coeff, double* diag) {
    if (!X || !in || !out || nd < 0 || in == out || in->v.ctx != &X->c || out->v.ctx != &X->c) return SFCI_E_ARG;
    Ctx* c = &X->c;
    return guarded(c, [&] {
        double d = 0.0;
        size_t n = c->screen_mode
                       ? dev_hv_screened(c, in->v.k, in->v.v, (size_t)in->v.len, (size_t)nd, lambda, coeff,
                                         (size_t)out->v.cap, out->v.k, out->v.v, &d)
                       : dev_hv(c, in->v.k, in->v.v, (size_t)in->v.len, (size_t)nd, lambda, coeff,
                                (size_t)out->v.cap, out->v.k, out->v.v, &d);
        out->v.len = (int64_t)n;
        if (diag) *diag = d;
    });
}

int sfci_set_screen(sfci_ctx* X, int mode, sfci_mlp* net, double keep_factor, double eps_pre) {
    if (!X || mode < SFCI_SCREEN_OFF || mode > SFCI_SCREEN_NETWORK_MAGNITUDE || (mode >= SFCI_SCREEN_NETWORK && !net) ||
        (mode != SFCI_SCREEN_OFF && (!(keep_factor > 0.0) || !(eps_pre >= 0.0))))
        return SFCI_E_ARG;
    Ctx* c = &X->c;
    c->screen_mode = mode;
    c->screen_net = mode >= SFCI_SCREEN_NETWORK ? net : nullptr;
    if (mode != SFCI_SCREEN_OFF) {
        c->screen_keep = keep_factor;
        c->screen_eps = eps_pre;
    }
    return SFCI_OK;
}

int sfci_get_screen(const sfci_ctx* X, int* mode, sfci_mlp** net, double* keep_factor, double* eps_pre) {
    if (!X) return SFCI_E_ARG;
    const Ctx* c = &X->c;
    if (mode) *mode = c->screen_mode;
    if (net) *net = const_cast<sfci_mlp*>(c->screen_net);
    if (keep_factor) *keep_factor = c->screen_keep;
    if (eps_pre) *eps_pre = c->screen_eps;
    return SFCI_OK;
}

int sfci_screen_samples(sfci_ctx* X, const sfci_vec* v, double eps_pre, int64_t max_samples, uint64_t seed,
                        sfci_screen_sample_t* out, int64_t* n_out) {
    if (!X || !v || !out || !n_out || max_samples < 0 || !(eps_pre >= 0.0) || v->v.ctx != &X->c) return SFCI_E_ARG;
    Ctx* c = &X->c;
    *n_out = 0;
    return guarded(c, [&] {
        *n_out = (int64_t)dev_screen_samples(c, v->v.k, v->v.v, (size_t)v->v.len, eps_pre, (size_t)max_samples, seed,
                                             out);
    });
}

int sfci_energy(sfci_ctx* X, const sfci_vec* v, double* e) {
    if (!X || !v || !e || v->v.ctx != &X->c) return SFCI_E_ARG;
    Ctx* c = &X->c;
    // FP64 throughout, whatever the H.v precision
    return guarded(c, [&] { *e = dev_energy(c, v->v.k, v->v.v, (size_t)v->v.len); });
}

int sfci_diag_element(sfci_ctx* X, uint64_t det, double* e) {
    if (!X || !e) return SFCI_E_ARG;
    Ctx* c = &X->c;
    if (popc64(det & 0x5555555555555555ull) != c->ham.k || popc64(det & 0xaaaaaaaaaaaaaaaaull) != c->ham.k ||
        (c->n_so < 64 && (det >> c->n_so))) {
        c->last_error = "determinant does not match n_so / n_elec";
        return SFCI_E_ARG;
    }
    return guarded(c, [&] {
        uint64_t* k = c->src_k.get<uint64_t>(1);
        double* v = c->src_v.get<double>(1);
        double one = 1.0;
        HIPCHK(hipMemcpyAsync(k, &det, 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(v, &one, 8, hipMemcpyHostToDevice, c->stream));
        // the input buffers are re-used by dev_hv only when n_in > nd, which is not the case here
        double d = 0.0;
        dev_hv(c, k, v, 1, 1, 0.0, 1.0, 0, nullptr, nullptr, &d);
        *e = d;
    });
}

int sfci_dot(sfci_ctx* X, const sfci_vec* a, const sfci_vec* b, double* out) {
    if (!X || !a || !b || !out) return SFCI_E_ARG;
    Ctx* c = &X->c;
    return guarded(c, [&] { *out = dev_dot(c, a->v.k, a->v.v, (size_t)a->v.len, b->v.k, b->v.v, (size_t)b->v.len); });
}

int sfci_norm(sfci_ctx* X, const sfci_vec* a, double* out) {
    if (!X || !a || !out) return SFCI_E_ARG;
    Ctx* c = &X->c;
    return guarded(c, [&] { *out = dev_norm(c, a->v.v, (size_t)a->v.len); });
}

int sfci_scale(sfci_ctx* X, sfci_vec* a, double s) {
    if (!X || !a) return SFCI_E_ARG;
    Ctx* c = &X->c;
    return guarded(c, [&] { dev_scale(c, a->v.v, (size_t)a->v.len, s); });
}

int sfci_add_two(sfci_ctx* X, sfci_vec* v, const sfci_vec* w, double a) {
    if (!X || !v || !w || v == w) return SFCI_E_ARG;
    Ctx* c = &X->c;
    return guarded(c, [&] {
        size_t n = dev_add_two(c, v->v.k, v->v.v, (size_t)v->v.len, w->v.k, w->v.v, (size_t)w->v.len, a,
                               (size_t)v->v.cap, v->v.k, v->v.v);
        v->v.len = (int64_t)n;
    });
}

int sfci_shrink_from(sfci_ctx* X, sfci_vec* dst, const sfci_vec* src) {
    if (!X || !dst || !src || dst == src) return SFCI_E_ARG;
    Ctx* c = &X->c;
    return guarded(c, [&] {
        size_t n = select_top(c, src->v.k, src->v.v, (size_t)src->v.len, (size_t)dst->v.cap, dst->v.k, dst->v.v);
        dst->v.len = (int64_t)n;
    });
}

int sfci_shrink_from_noalloc(sfci_ctx* X, sfci_vec* dst, const sfci_vec* src) {
    if (!X || !dst || !src || dst == src) return SFCI_E_ARG;
    Ctx* c = &X->c;
    return guarded(c, [&] {
        size_t n = dev_shrink_noalloc(c, src->v.k, src->v.v, (size_t)src->v.len, (size_t)dst->v.cap, dst->v.k,
                                      dst->v.v);
        dst->v.len = (int64_t)n;
    });
}

int sfci_lincomb(sfci_ctx* X, const sfci_vec* const* vecs, const double* coef, int n, sfci_vec* out) {
    if (!X || !out || n < 0 || (n && (!vecs || !coef))) return SFCI_E_ARG;
    Ctx* c = &X->c;
    std::vector<const Vec*> vs(n);
    for (int i = 0; i < n; ++i) {
        if (!vecs[i]) return SFCI_E_ARG;
        vs[i] = &vecs[i]->v;
    }
    return guarded(c, [&] {
        size_t m = dev_lincomb(c, vs.data(), coef, n, (size_t)out->v.cap, out->v.k, out->v.v);
        out->v.len = (int64_t)m;
    });
}

// Cut-then-cap tail of the linear-solver products (src/functions.cpp:1016-1062):
// dst = top_{dst.cap}({|c| > cut} of src), key order.
int sfci_shrink_cut(sfci_ctx* X, sfci_vec* dst, const sfci_vec* src, double cut) {
    if (!X || !dst || !src || dst == src) return SFCI_E_ARG;
    Ctx* c = &X->c;
    return guarded(c, [&] {
        size_t n = (size_t)src->v.len;
        uint64_t* fk = c->lin_k[0].get<uint64_t>(std::max<size_t>(n, 1));
        double* fv = c->lin_v[0].get<double>(std::max<size_t>(n, 1));
        size_t f = dev_filter_above(c, src->v.k, src->v.v, n, cut, fk, fv);
        dst->v.len = (int64_t)select_top(c, fk, fv, f, (size_t)dst->v.cap, dst->v.k, dst->v.v);
    });
}

namespace {
// The uncut product w = coeff (H - lambda) in, copied into lin_k[0]/lin_v[0] with room for
// `extra` more entries; returns its length.
size_t hv_all_into_lin0(Ctx* c, const Vec& in, double lambda, size_t extra, size_t* cap_out) {
    double d = 0.0;
    HvAll a;
    dev_hv_all(c, in.k, in.v, (size_t)in.len, lambda, 1.0, &d, &a);
    size_t cap = a.n + extra + 1;
    uint64_t* k0 = c->lin_k[0].get<uint64_t>(cap);
    double* v0 = c->lin_v[0].get<double>(cap);
    c->lin_k[1].get<uint64_t>(cap);
    c->lin_v[1].get<double>(cap);
    dev_copy(c, k0, v0, a.k, a.v, a.n);
    *cap_out = cap;
    return a.n;
}
}  // namespace

// spmspvm_bucket_linear (src/functions.cpp:858-1094): w = (H - lambda) in, uncut;
// hij[j] = <basis[j] | w> for every j < n_basis; w -= hij[n-1] basis[n-1], then
// w -= hij[n-2] basis[n-2]; out = top_{out.cap}({|w| > cut}), key order.
int sfci_hv_linear(sfci_ctx* X, const sfci_vec* in, sfci_vec* out, double lambda, const sfci_vec* const* basis,
                   int n_basis, double* hij) {
    if (!X || !in || !out || !basis || n_basis < 1 || !hij || out == in) return SFCI_E_ARG;
    for (int j = 0; j < n_basis; ++j)
        if (!basis[j] || basis[j] == out) return SFCI_E_ARG;
    Ctx* c = &X->c;
    return guarded(c, [&] {
        for (int j = 0; j < n_basis; ++j) hij[j] = 0.0;
        out->v.len = 0;
        if (in->v.len == 0) return;
        const Vec& bc = basis[n_basis - 1]->v;
        const Vec* bp = n_basis > 1 ? &basis[n_basis - 2]->v : nullptr;
        size_t cap = 0;
        size_t m = hv_all_into_lin0(c, in->v, lambda, (size_t)bc.len + (bp ? (size_t)bp->len : 0), &cap);
        uint64_t* k[2] = {(uint64_t*)c->lin_k[0].p, (uint64_t*)c->lin_k[1].p};
        double* v[2] = {(double*)c->lin_v[0].p, (double*)c->lin_v[1].p};
        for (int j = 0; j < n_basis; ++j)
            hij[j] = dev_dot(c, k[0], v[0], m, basis[j]->v.k, basis[j]->v.v, (size_t)basis[j]->v.len);
        m = dev_add_two(c, k[0], v[0], m, bc.k, bc.v, (size_t)bc.len, -hij[n_basis - 1], cap, k[1], v[1]);
        int cur = 1;
        if (bp) {
            m = dev_add_two(c, k[1], v[1], m, bp->k, bp->v, (size_t)bp->len, -hij[n_basis - 2], cap, k[0], v[0]);
            cur = 0;
        }
        size_t f = dev_filter_above(c, k[cur], v[cur], m, c->cut, k[1 - cur], v[1 - cur]);
        out->v.len = (int64_t)select_top(c, k[1 - cur], v[1 - cur], f, (size_t)out->v.cap, out->v.k, out->v.v);
    });
}

// spmspvm_bucket_res_norm (src/functions.cpp:1096-1259): w = (H - lambda) x - b, uncut;
// *res = ||w||; out = top_{out.cap}({|w| > cut}), key order.
int sfci_hv_residual(sfci_ctx* X, const sfci_vec* x, sfci_vec* out, double lambda, const sfci_vec* b, double* res) {
    if (!X || !x || !out || !b || !res || out == x || out == b) return SFCI_E_ARG;
    Ctx* c = &X->c;
    return guarded(c, [&] {
        size_t cap = 0;
        size_t m = hv_all_into_lin0(c, x->v, lambda, (size_t)b->v.len, &cap);
        uint64_t* k[2] = {(uint64_t*)c->lin_k[0].p, (uint64_t*)c->lin_k[1].p};
        double* v[2] = {(double*)c->lin_v[0].p, (double*)c->lin_v[1].p};
        m = dev_add_two(c, k[0], v[0], m, b->v.k, b->v.v, (size_t)b->v.len, -1.0, cap, k[1], v[1]);
        *res = dev_norm(c, v[1], m);
        size_t f = dev_filter_above(c, k[1], v[1], m, c->cut, k[0], v[0]);
        out->v.len = (int64_t)select_top(c, k[0], v[0], f, (size_t)out->v.cap, out->v.k, out->v.v);
    });
}

// calc_CISD (src/sparse_solver.cpp:15-114)
int sfci_cisd(sfci_ctx* X, uint64_t ref_det, int state_idx, sfci_vec* out, double* eigenvalue, int64_t* block_size) {
    if (!X || !out || !eigenvalue || state_idx < 0) return SFCI_E_ARG;
    Ctx* c = &X->c;
    if (popc64(ref_det & 0x5555555555555555ull) != c->ham.k || popc64(ref_det & 0xaaaaaaaaaaaaaaaaull) != c->ham.k ||
        (c->n_so < 64 && (ref_det >> c->n_so))) {
        c->last_error = "determinant does not match n_so / n_elec";
        return SFCI_E_ARG;
    }
    return guarded(c, [&] {
        size_t n = dev_cisd(c, ref_det, state_idx, (size_t)out->v.cap, out->v.k, out->v.v, eigenvalue);
        out->v.len = (int64_t)n;
        if (block_size) *block_size = (int64_t)n;
    });
}

// dynVec::addThree (src/all.h:482-608): out = top_{out.cap}(a x + b y + c z), then the
// entries with |c| <= 1e-14 dropped, key order.
int sfci_add_three(sfci_ctx* X, sfci_vec* out, const sfci_vec* x, const sfci_vec* y, const sfci_vec* z, double a,
                   double b, double cz) {
    if (!X || !out || !x || !y || !z) return SFCI_E_ARG;
    Ctx* c = &X->c;
    return guarded(c, [&] {
        const Vec* vs[3] = {&x->v, &y->v, &z->v};
        const double co[3] = {a, b, cz};
        size_t cap = std::max<size_t>((size_t)out->v.cap, 1);
        uint64_t* k0 = c->lin_k[0].get<uint64_t>(cap);
        double* v0 = c->lin_v[0].get<double>(cap);
        uint64_t* k1 = c->lin_k[1].get<uint64_t>(cap);
        double* v1 = c->lin_v[1].get<double>(cap);
        size_t m = dev_lincomb(c, vs, co, 3, (size_t)out->v.cap, k0, v0);
        size_t f = dev_filter_above(c, k0, v0, m, 1e-14, k1, v1);
        dev_copy(c, out->v.k, out->v.v, k1, v1, f);
        out->v.len = (int64_t)f;
    });
}

// ---- 128-bit keys (more than 64 spin orbitals) -------------------------------------

int sfci_wide_create(const char* in_path, int n_so, int n_elec, int device, sfci_wctx** out) {
    if (!in_path || !out) return SFCI_E_ARG;
    *out = nullptr;
    if (n_elec >= 64) {
        // the merge marks empty LDS-table slots with an all-ones key word, which a
        // determinant with fewer than 64 electrons never has
        fprintf(stderr, "sfci_wide_create: n_elec must be below 64\n");
        return SFCI_E_ARG;
    }
    std::vector<Entry> ent;
    int rc = read_in_file(in_path, ent);
    if (rc) return rc;
    std::string err;
    sfci_ctx* X = nullptr;
    rc = build_ctx(n_so, n_elec, ent, device, &X, err, 128);
    if (rc) {
        fprintf(stderr, "sfci_wide_create: %s\n", err.c_str());
        return rc;
    }
    *out = new sfci_wctx{X};
    return SFCI_OK;
}

void sfci_wide_destroy(sfci_wctx* W) {
    if (!W) return;
    sfci_destroy(W->x);
    delete W;
}

const char* sfci_wide_last_error(const sfci_wctx* W) { return W ? sfci_last_error(W->x) : "null context"; }

int sfci_wide_info(const sfci_wctx* W, sfci_info_t* info) { return W ? sfci_info(W->x, info) : SFCI_E_ARG; }

int sfci_wide_set_cut(sfci_wctx* W, double cut) { return W ? sfci_set_cut(W->x, cut) : SFCI_E_ARG; }

int sfci_wide_set_precision(sfci_wctx* W, int precision) {
    return W ? sfci_set_precision(W->x, precision) : SFCI_E_ARG;
}

int sfci_wide_set_workspace(sfci_wctx* W, uint64_t bytes) { return W ? sfci_set_workspace(W->x, bytes) : SFCI_E_ARG; }

int sfci_wide_set_merge(sfci_wctx* W, int hash_bits, int table_fill) {
    if (!W || !(hash_bits == 0 || (hash_bits >= 9 && hash_bits <= 16)) || table_fill < 0 || table_fill > 2867)
        return SFCI_E_ARG;
    W->x->c.wide_merge_bits = hash_bits;
    W->x->c.wide_fill = table_fill;
    return SFCI_OK;
}

int sfci_wide_hv_stats(const sfci_wctx* W, sfci_hv_stats_t* s) { return W ? sfci_hv_stats(W->x, s) : SFCI_E_ARG; }

int sfci_wide_vec_create(sfci_wctx* W, int64_t cap, sfci_wvec** out) {
    if (!W || !out || cap < 1) return SFCI_E_ARG;
    *out = nullptr;
    Ctx* c = &W->x->c;
    sfci_wvec* V = new sfci_wvec();
    int rc = guarded(c, [&] {
        V->ctx = c;
        V->cap = cap;
        HIPCHK(hipMalloc(&V->k, (size_t)cap * 16));
        HIPCHK(hipMalloc(&V->v, (size_t)cap * 8));
    });
    if (rc) {
        if (V->k) (void)hipFree(V->k);
        if (V->v) (void)hipFree(V->v);
        delete V;
        return rc;
    }
    *out = V;
    return SFCI_OK;
}

void sfci_wide_vec_destroy(sfci_wvec* V) {
    if (!V) return;
    (void)hipSetDevice(V->ctx->device);
    (void)hipStreamSynchronize(V->ctx->stream);
    if (V->k) (void)hipFree(V->k);
    if (V->v) (void)hipFree(V->v);
    delete V;
}

int64_t sfci_wide_vec_len(const sfci_wvec* V) { return V ? V->len : -1; }

int sfci_wide_vec_upload(sfci_wvec* V, const uint64_t* keys, const double* vals, int64_t n) {
    if (!V || n < 0 || n > V->cap || (n && (!keys || !vals))) return SFCI_E_ARG;
    Ctx* c = V->ctx;
    const uint64_t ev = 0x5555555555555555ull, od = 0xaaaaaaaaaaaaaaaaull;
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t blo = keys[2 * i], bhi = keys[2 * i + 1];
        if (i > 0) {
            const uint64_t alo = keys[2 * i - 2], ahi = keys[2 * i - 1];
            if (!(ahi < bhi || (ahi == bhi && alo < blo))) {
                set_err(c, "sfci_wide_vec_upload: keys must be strictly ascending (hi:lo)");
                return SFCI_E_ARG;
            }
        }
        const bool beyond = c->n_so <= 64 ? (bhi != 0 || (c->n_so < 64 && (blo >> c->n_so)))
                                          : (c->n_so < 128 && (bhi >> (c->n_so - 64)));
        if (popc64(blo & ev) + popc64(bhi & ev) != c->ham.k || popc64(blo & od) + popc64(bhi & od) != c->ham.k ||
            beyond) {
            set_err(c, "sfci_wide_vec_upload: key " + std::to_string(i) + " does not match n_so / n_elec");
            return SFCI_E_ARG;
        }
    }
    return guarded(c, [&] {
        if (n) {
            HIPCHK(hipMemcpyAsync(V->k, keys, (size_t)n * 16, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(V->v, vals, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        V->len = n;
    });
}

int sfci_wide_vec_download(const sfci_wvec* V, uint64_t* keys, double* vals, int64_t max_n, int64_t* n_out) {
    if (!V || max_n < 0) return SFCI_E_ARG;
    const int64_t n = std::min(V->len, max_n);
    if (n && (!keys || !vals)) return SFCI_E_ARG;
    Ctx* c = V->ctx;
    return guarded(c, [&] {
        if (n) {
            HIPCHK(hipMemcpyAsync(keys, V->k, (size_t)n * 16, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(vals, V->v, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        if (n_out) *n_out = n;
    });
}

int sfci_wide_hv(sfci_wctx* W, const sfci_wvec* in, sfci_wvec* out, int64_t nd, double lambda, double coeff,
                 double* diag) {
    if (!W || !in || !out || nd < 0 || in == out || in->ctx != &W->x->c || out->ctx != &W->x->c) return SFCI_E_ARG;
    Ctx* c = &W->x->c;
    return guarded(c, [&] {
        double d = 0.0;
        const size_t n = dev_hv_wide(c, c->n_so, in->k, in->v, (size_t)in->len, (size_t)nd, lambda, coeff,
                                     (size_t)out->cap, out->k, out->v, &d);
        out->len = (int64_t)n;
        if (diag) *diag = d;
    });
}

int sfci_wide_vec_copy(sfci_wvec* D, const sfci_wvec* S) {
    if (!D || !S || D->ctx != S->ctx || S->len > D->cap) return SFCI_E_ARG;
    if (D == S) return SFCI_OK;
    Ctx* c = D->ctx;
    return guarded(c, [&] {
        if (S->len) {
            HIPCHK(hipMemcpyAsync(D->k, S->k, (size_t)S->len * 16, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(D->v, S->v, (size_t)S->len * 8, hipMemcpyDeviceToDevice, c->stream));
        }
        D->len = S->len;
    });
}

int sfci_wide_vec_clear(sfci_wvec* V) {
    if (!V) return SFCI_E_ARG;
    V->len = 0;
    return SFCI_OK;
}

int sfci_wide_dot(sfci_wctx* W, const sfci_wvec* a, const sfci_wvec* b, double* out) {
    if (!W || !a || !b || !out || a->ctx != &W->x->c || b->ctx != &W->x->c) return SFCI_E_ARG;
    Ctx* c = &W->x->c;
    return guarded(c, [&] { *out = dev_wide_dot(c, a->k, a->v, (size_t)a->len, b->k, b->v, (size_t)b->len); });
}

int sfci_wide_norm(sfci_wctx* W, const sfci_wvec* a, double* out) {
    if (!W || !a || !out || a->ctx != &W->x->c) return SFCI_E_ARG;
    Ctx* c = &W->x->c;
    return guarded(c, [&] { *out = dev_norm(c, a->v, (size_t)a->len); });
}

int sfci_wide_scale(sfci_wctx* W, sfci_wvec* a, double s) {
    if (!W || !a || a->ctx != &W->x->c) return SFCI_E_ARG;
    Ctx* c = &W->x->c;
    return guarded(c, [&] { dev_scale(c, a->v, (size_t)a->len, s); });
}

int sfci_wide_add_two(sfci_wctx* W, sfci_wvec* v, const sfci_wvec* w, double a) {
    if (!W || !v || !w || v == w || v->ctx != &W->x->c || w->ctx != &W->x->c) return SFCI_E_ARG;
    Ctx* c = &W->x->c;
    return guarded(c, [&] {
        v->len = (int64_t)dev_wide_add_two(c, v->k, v->v, (size_t)v->len, w->k, w->v, (size_t)w->len, a,
                                           (size_t)v->cap, v->k, v->v);
    });
}

int sfci_wide_shrink_from(sfci_wctx* W, sfci_wvec* dst, const sfci_wvec* src) {
    if (!W || !dst || !src || dst->ctx != &W->x->c || src->ctx != &W->x->c) return SFCI_E_ARG;
    Ctx* c = &W->x->c;
    return guarded(c, [&] {
        dst->len = (int64_t)dev_wide_shrink(c, src->k, src->v, (size_t)src->len, (size_t)dst->cap, dst->k, dst->v);
    });
}

int sfci_wide_lincomb(sfci_wctx* W, const sfci_wvec* const* vecs, const double* coef, int n, sfci_wvec* out) {
    if (!W || !vecs || !coef || n < 1 || !out || out->ctx != &W->x->c) return SFCI_E_ARG;
    for (int i = 0; i < n; ++i)
        if (!vecs[i] || vecs[i]->ctx != &W->x->c) return SFCI_E_ARG;
    Ctx* c = &W->x->c;
    return guarded(c, [&] {
        out->len = (int64_t)dev_wide_lincomb(c, vecs, coef, n, (size_t)out->cap, out->k, out->v);
    });
}

int sfci_wide_energy(sfci_wctx* W, const sfci_wvec* v, double* e) {
    if (!W || !v || !e || v->ctx != &W->x->c) return SFCI_E_ARG;
    Ctx* c = &W->x->c;
    return guarded(c, [&] { *e = dev_wide_energy(c, v->k, v->v, (size_t)v->len); });
}

int sfci_wide_sync(sfci_wctx* W) { return W ? sfci_sync(W->x) : SFCI_E_ARG; }

}  // extern "C"
