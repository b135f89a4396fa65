// Sparse-vector operations of the Lanczos recurrences (src/all.h:249-480) on the GPU.
//
//   select_top  top-cap by |c| (ties: smaller key first) -- the |amp| introsort plus
//               "keep the last cap" of addTwo/shrinkFrom/SpMVMtply_bucket, done as an
//               MSD radix select on the |c| bit pattern followed by a stable compaction
//   dot         merge intersection (src/all.h:317-351), exact accumulation
//   norm/scale  src/all.h:249-271
//   add_two     v + a*w merged by key, zeros kept, then top-cap (src/all.h:407-480)
//   lincomb     restart recombination sum_i c_i K_i (src/sparse_solver.cpp:492-509)
#include "internal.h"

namespace sfci {

static inline int grid_of(size_t n, int cap = 256 * 16) {
    size_t g = (n + SFCI_BLOCK - 1) / SFCI_BLOCK;
    if (g < 1) g = 1;
    if (g > (size_t)cap) g = cap;
    return (int)g;
}

__device__ __forceinline__ uint64_t abs_bits(double x) {
    uint64_t b;
    memcpy(&b, &x, 8);
    return b & 0x7fffffffffffffffull;
}

// ---- kernel timing -----------------------------------------------------------------
static hipEvent_t kt_event(Ctx* c) {
    auto& k = c->kclock;
    if (k.used == k.pool.size()) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        k.pool.push_back(e);
    }
    return k.pool[k.used++];
}

int kt_begin(Ctx* c, int id, double bytes) {
    auto& k = c->kclock;
    if (k.recs.size() >= 4096) kt_resolve(c);
    KernelClock::Rec r{id, bytes, kt_event(c), nullptr};
    HIPCHK(hipEventRecord(r.a, c->stream));
    k.recs.push_back(r);
    return (int)k.recs.size() - 1;
}

void kt_end(Ctx* c, int rec) {
    auto& r = c->kclock.recs[rec];
    r.b = kt_event(c);
    HIPCHK(hipEventRecord(r.b, c->stream));
}

void kt_resolve(Ctx* c) {
    auto& k = c->kclock;
    if (!k.recs.empty() || !k.phases.empty()) {
        HIPCHK(hipStreamSynchronize(c->stream));
        for (auto& r : k.recs) {
            if (!r.b) continue;
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
            k.stat[r.id].ms += ms;
            k.stat[r.id].launches += 1;
            k.stat[r.id].bytes += r.bytes;
        }
        for (auto& p : k.phases) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
            *p.dst += ms;
        }
        k.recs.clear();
        k.phases.clear();
    }
    // an open phase timer still holds its start event: the pool is reused only when none is
    if (k.open == 0) k.used = 0;
}

PhaseTimer::PhaseTimer(Ctx* ctx) : c(ctx) {
    a = kt_event(c);
    ++c->kclock.open;
    HIPCHK(hipEventRecord(a, c->stream));
}

PhaseTimer::~PhaseTimer() {
    auto& k = c->kclock;
    if (--k.open == 0 && k.recs.empty() && k.phases.empty()) k.used = 0;
}

void PhaseTimer::lap(double* dst) {
    hipEvent_t b = kt_event(c);
    HIPCHK(hipEventRecord(b, c->stream));
    c->kclock.phases.push_back(KernelClock::Phase{dst, a, b});
    a = b;
}

// ---- accumulators ------------------------------------------------------------------
void acc_clear(Ctx* c, int slot) {
    HIPCHK(hipMemsetAsync(c->d_acc + 2 * slot, 0, 16, c->stream));
}

fx128 acc_read_raw(Ctx* c, int slot) {
    uint64_t w[2];
    HIPCHK(hipMemcpyAsync(w, c->d_acc + 2 * slot, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return fx128{w[0], w[1]};
}

double acc_read(Ctx* c, int slot) { return fx_to(acc_read_raw(c, slot)); }

uint32_t flags_read_clear(Ctx* c) {
    uint32_t f = 0;
    HIPCHK(hipMemcpyAsync(&f, c->d_flags, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemsetAsync(c->d_flags, 0, 4, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return f;
}

void atomic_flags_restore(Ctx* c, uint32_t f) {
    if (!f) return;
    uint32_t cur = 0;
    HIPCHK(hipMemcpyAsync(&cur, c->d_flags, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    cur |= f;
    HIPCHK(hipMemcpyAsync(c->d_flags, &cur, 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
}

// ---- elementwise -------------------------------------------------------------------
__global__ void k_copy(uint64_t* __restrict__ dk, double* __restrict__ dv, const uint64_t* __restrict__ sk,
                       const double* __restrict__ sv, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        dk[i] = sk[i];
        dv[i] = sv[i];
    }
}

void dev_copy(Ctx* c, uint64_t* dk, double* dv, const uint64_t* sk, const double* sv, size_t n) {
    if (n == 0 || dk == sk) return;
    k_copy<<<grid_of(n), SFCI_BLOCK, 0, c->stream>>>(dk, dv, sk, sv, n);
    HIPCHK(hipGetLastError());
}

__global__ void k_scale(double* __restrict__ v, size_t n, double s) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v[i] = v[i] * s;
}

void dev_scale(Ctx* c, double* v, size_t n, double s) {
    if (n == 0) return;
    k_scale<<<grid_of(n), SFCI_BLOCK, 0, c->stream>>>(v, n, s);
    HIPCHK(hipGetLastError());
}

__global__ __launch_bounds__(SFCI_BLOCK) void k_norm2(const double* __restrict__ v, size_t n,
                                                      unsigned long long* acc, uint32_t* flags) {
    fx128 a = fx_zero();
    bool bad = false;
    for (size_t i = (size_t)blockIdx.x * SFCI_BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * SFCI_BLOCK) {
        double p = v[i] * v[i];
        bad |= !fx_in_range(p);
        a = fx_add(a, fx_from(p));
    }
    if (bad) atomicOr(flags, FLAG_RANGE);
    fx_block_reduce_atomic<SFCI_BLOCK>(a, acc);
}

double dev_norm(Ctx* c, const double* v, size_t n) {
    acc_clear(c, 2);
    if (n) k_norm2<<<grid_of(n), SFCI_BLOCK, 0, c->stream>>>(v, n, c->d_acc + 4, c->d_flags);
    HIPCHK(hipGetLastError());
    double s = acc_read(c, 2);
    if (flags_read_clear(c) & FLAG_RANGE) throw HipError("norm: term outside the fixed-point range", -5);
    return sqrt(s);
}

// a is searched element-wise in b (b sorted by key)
__global__ __launch_bounds__(SFCI_BLOCK) void k_dot(const uint64_t* __restrict__ ak, const double* __restrict__ av,
                                                    size_t na, const uint64_t* __restrict__ bk,
                                                    const double* __restrict__ bv, size_t nb,
                                                    unsigned long long* acc, uint32_t* flags) {
    fx128 a = fx_zero();
    bool bad = false;
    for (size_t i = (size_t)blockIdx.x * SFCI_BLOCK + threadIdx.x; i < na; i += (size_t)gridDim.x * SFCI_BLOCK) {
        uint64_t key = ak[i];
        size_t lo = 0, hi = nb;
        while (lo < hi) {
            size_t mid = (lo + hi) >> 1;
            if (bk[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        if (lo < nb && bk[lo] == key) {
            double p = av[i] * bv[lo];
            bad |= !fx_in_range(p);
            a = fx_add(a, fx_from(p));
        }
    }
    if (bad) atomicOr(flags, FLAG_RANGE);
    fx_block_reduce_atomic<SFCI_BLOCK>(a, acc);
}

double dev_dot(Ctx* c, const uint64_t* ak, const double* av, size_t na, const uint64_t* bk, const double* bv,
               size_t nb) {
    acc_clear(c, 3);
    if (na && nb) {
        if (na <= nb) k_dot<<<grid_of(na), SFCI_BLOCK, 0, c->stream>>>(ak, av, na, bk, bv, nb, c->d_acc + 6, c->d_flags);
        else k_dot<<<grid_of(nb), SFCI_BLOCK, 0, c->stream>>>(bk, bv, nb, ak, av, na, c->d_acc + 6, c->d_flags);
        HIPCHK(hipGetLastError());
    }
    double d = acc_read(c, 3);
    if (flags_read_clear(c) & FLAG_RANGE) throw HipError("dot: term outside the fixed-point range", -5);
    return d;
}

// ---- top-cap selection -------------------------------------------------------------
// histogram of digit (|v| >> shift) & 255 among items whose |v| bits match prefix under mask
__global__ __launch_bounds__(SFCI_BLOCK) void k_sel_hist(const double* __restrict__ v, size_t n, uint64_t prefix,
                                                         uint64_t mask, int shift, unsigned long long* hist) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * SFCI_BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * SFCI_BLOCK) {
        uint64_t b = abs_bits(v[i]);
        if ((b & mask) == prefix) atomicAdd(&h[(b >> shift) & 255], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// collect keys of items with |v| bits == tau
__global__ void k_sel_ties(const uint64_t* __restrict__ k, const double* __restrict__ v, size_t n, uint64_t tau,
                           uint64_t* __restrict__ tk, unsigned long long* count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (abs_bits(v[i]) == tau) {
            unsigned long long p = atomicAdd(count, 1ull);
            tk[p] = k[i];
        }
    }
}

__global__ void k_sel_flags(const uint64_t* __restrict__ k, const double* __restrict__ v, size_t n, uint64_t tau,
                            uint64_t key_cut, uint32_t* __restrict__ fl) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint64_t b = abs_bits(v[i]);
        fl[i] = (b > tau || (b == tau && k[i] <= key_cut)) ? 1u : 0u;
    }
}

__global__ void k_compact(const uint64_t* __restrict__ k, const double* __restrict__ v, size_t n,
                          const uint32_t* __restrict__ fl, const uint64_t* __restrict__ pos,
                          uint64_t* __restrict__ ok, double* __restrict__ ov) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (fl[i]) {
            ok[pos[i]] = k[i];
            ov[pos[i]] = v[i];
        }
    }
}

size_t select_top(Ctx* c, const uint64_t* k, const double* v, size_t n, size_t cap, uint64_t* ok, double* ov) {
    if (n <= cap) {
        dev_copy(c, ok, ov, k, v, n);
        return n;
    }
    if (cap == 0) return 0;
    unsigned long long* hist = (unsigned long long*)c->scratch_small.get<uint64_t>(256 + 8);
    std::vector<unsigned long long> h(256);
    uint64_t prefix = 0, mask = 0;
    uint64_t need = cap;   // how many still to take among items matching prefix
    for (int shift = 56; shift >= 0; shift -= 8) {
        HIPCHK(hipMemsetAsync(hist, 0, 256 * sizeof(unsigned long long), c->stream));
        k_sel_hist<<<grid_of(n), SFCI_BLOCK, 0, c->stream>>>(v, n, prefix, mask, shift, hist);
        HIPCHK(hipMemcpyAsync(h.data(), hist, 256 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        uint64_t above = 0;
        int d = 255;
        for (; d >= 0; --d) {
            if (above + h[d] >= need) break;
            above += h[d];
        }
        if (d < 0) throw HipError("select_top: inconsistent histogram", -1);
        need -= above;
        prefix |= (uint64_t)d << shift;
        mask |= (uint64_t)255 << shift;
    }
    const uint64_t tau = prefix;
    // need = how many of the items equal to tau are kept; smallest keys first
    uint64_t key_cut = ~0ull;
    {
        unsigned long long* cnt = (unsigned long long*)c->d_counter + 4;
        HIPCHK(hipMemsetAsync(cnt, 0, 8, c->stream));
        // the number of ties is at most n - (cap - need)
        uint64_t* tk = c->sel_a.get<uint64_t>(n - (cap - need) + 1);
        k_sel_ties<<<grid_of(n), SFCI_BLOCK, 0, c->stream>>>(k, v, n, tau, tk, cnt);
        unsigned long long nt = 0;
        HIPCHK(hipMemcpyAsync(&nt, cnt, 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (nt > need) {
            // sort the tied keys (values unused)
            double* tv = c->sel_b.get<double>(nt);
            uint64_t* tk2 = c->sel_c.get<uint64_t>(2 * nt);
            double* tv2 = (double*)(tk2 + nt);
            HIPCHK(hipMemsetAsync(tv, 0, nt * sizeof(double), c->stream));
            uint64_t* kk = tk;
            double* vv = tv;
            radix_sort_pairs(c, &kk, &vv, tk2, tv2, nt, 64);
            HIPCHK(hipMemcpyAsync(&key_cut, kk + (need - 1), 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
    }
    uint32_t* fl = c->sel_b.get<uint32_t>(n);
    k_sel_flags<<<grid_of(n), SFCI_BLOCK, 0, c->stream>>>(k, v, n, tau, key_cut, fl);
    uint64_t* pos = c->sel_c.get<uint64_t>(n + 1);
    exclusive_scan_u32(c, fl, n, pos, pos + n);
    k_compact<<<grid_of(n), SFCI_BLOCK, 0, c->stream>>>(k, v, n, fl, pos, ok, ov);
    HIPCHK(hipGetLastError());
    return cap;
}

// ---- magnitude filter --------------------------------------------------------------
__global__ void k_flag_above(const double* __restrict__ v, size_t n, double cut, uint32_t* __restrict__ fl) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        fl[i] = fabs(v[i]) > cut ? 1u : 0u;
}

size_t dev_filter_above(Ctx* c, const uint64_t* k, const double* v, size_t n, double cut, uint64_t* ok, double* ov) {
    if (n == 0) return 0;
    uint32_t* fl = c->sel_b.get<uint32_t>(n);
    uint64_t* pos = c->sel_c.get<uint64_t>(n + 1);
    k_flag_above<<<grid_of(n), SFCI_BLOCK, 0, c->stream>>>(v, n, cut, fl);
    exclusive_scan_u32(c, fl, n, pos, pos + n);
    k_compact<<<grid_of(n), SFCI_BLOCK, 0, c->stream>>>(k, v, n, fl, pos, ok, ov);
    HIPCHK(hipGetLastError());
    uint64_t m = 0;
    HIPCHK(hipMemcpyAsync(&m, pos + n, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return m;
}

// ---- add_two -----------------------------------------------------------------------
// For each element of x, the lower bound in y and whether the key is present in y.
__global__ void k_lower_bound(const uint64_t* __restrict__ xk, size_t nx, const uint64_t* __restrict__ yk, size_t ny,
                              uint64_t* __restrict__ lb, uint32_t* __restrict__ dup) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nx; i += (size_t)gridDim.x * blockDim.x) {
        uint64_t key = xk[i];
        size_t lo = 0, hi = ny;
        while (lo < hi) {
            size_t mid = (lo + hi) >> 1;
            if (yk[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        lb[i] = lo;
        dup[i] = (lo < ny && yk[lo] == key) ? 1u : 0u;
    }
}

// merged[i + lb_w[i] - cdup_v[i]] = v_i (+ a*w if duplicated)
__global__ void k_merge_v(const uint64_t* __restrict__ vk, const double* __restrict__ vv, size_t nv,
                          const double* __restrict__ wv, double a, const uint64_t* __restrict__ lbw,
                          const uint32_t* __restrict__ dupv, const uint64_t* __restrict__ cdv,
                          uint64_t* __restrict__ ok, double* __restrict__ ov) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (size_t)gridDim.x * blockDim.x) {
        size_t p = i + lbw[i] - cdv[i];
        ok[p] = vk[i];
        ov[p] = dupv[i] ? (a * wv[lbw[i]] + vv[i]) : vv[i];
    }
}

__global__ void k_merge_w(const uint64_t* __restrict__ wk, const double* __restrict__ wv, size_t nw, double a,
                          const uint64_t* __restrict__ lbv, const uint32_t* __restrict__ dupw,
                          const uint64_t* __restrict__ cdw, uint64_t* __restrict__ ok, double* __restrict__ ov) {
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < nw; j += (size_t)gridDim.x * blockDim.x) {
        if (dupw[j]) continue;
        size_t p = j + lbv[j] - cdw[j];
        ok[p] = wk[j];
        ov[p] = a * wv[j];
    }
}

size_t dev_add_two(Ctx* c, const uint64_t* vk, const double* vv, size_t nv, const uint64_t* wk, const double* wv,
                   size_t nw, double a, size_t cap, uint64_t* ok, double* ov) {
    // scratch: lb/dup/prefix for both sides, then the merged list
    size_t nmax = nv + nw + 2;
    uint64_t* lbw = c->sel_a.get<uint64_t>(3 * nmax);
    uint64_t* lbv = lbw + nmax;
    uint64_t* cdv = lbv + nmax;
    uint64_t* cdw = c->sel_c.get<uint64_t>(nmax);
    uint32_t* dupv = c->sel_b.get<uint32_t>(2 * nmax);
    uint32_t* dupw = dupv + nmax;
    if (nv) k_lower_bound<<<grid_of(nv), SFCI_BLOCK, 0, c->stream>>>(vk, nv, wk, nw, lbw, dupv);
    if (nw) k_lower_bound<<<grid_of(nw), SFCI_BLOCK, 0, c->stream>>>(wk, nw, vk, nv, lbv, dupw);
    exclusive_scan_u32(c, dupv, nv, cdv, cdv + nv);
    exclusive_scan_u32(c, dupw, nw, cdw, nullptr);
    uint64_t ndup = 0;
    HIPCHK(hipMemcpyAsync(&ndup, cdv + nv, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    size_t nm = nv + nw - ndup;
    uint64_t* mk = c->tmp_k[0].get<uint64_t>(nm);
    double* mv = c->tmp_v[0].get<double>(nm);
    if (nv) k_merge_v<<<grid_of(nv), SFCI_BLOCK, 0, c->stream>>>(vk, vv, nv, wv, a, lbw, dupv, cdv, mk, mv);
    if (nw) k_merge_w<<<grid_of(nw), SFCI_BLOCK, 0, c->stream>>>(wk, wv, nw, a, lbv, dupw, cdw, mk, mv);
    HIPCHK(hipGetLastError());
    return select_top(c, mk, mv, nm, cap, ok, ov);
}

// ---- lincomb -----------------------------------------------------------------------
__global__ void k_scaled_copy(const uint64_t* __restrict__ sk, const double* __restrict__ sv, size_t n, double a,
                              uint64_t* __restrict__ dk, double* __restrict__ dv) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        dk[i] = sk[i];
        dv[i] = a * sv[i];
    }
}

// sorted (key, value) -> unique keys with exact sums; stable compaction via scan
__global__ void k_seg_heads(const uint64_t* __restrict__ k, size_t n, uint32_t* __restrict__ head) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        head[i] = (i == 0 || k[i - 1] != k[i]) ? 1u : 0u;
}

__global__ void k_seg_sum(const uint64_t* __restrict__ k, const double* __restrict__ v, size_t n,
                          const uint32_t* __restrict__ head, const uint64_t* __restrict__ pos,
                          uint64_t* __restrict__ ok, double* __restrict__ ov, uint32_t* flags) {
    bool bad = false;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (!head[i]) continue;
        fx128 acc = fx_zero();
        size_t j = i;
        while (j < n && k[j] == k[i]) {
            bad |= !fx_in_range(v[j]);
            acc = fx_add(acc, fx_from(v[j]));
            ++j;
        }
        ok[pos[i]] = k[i];
        ov[pos[i]] = fx_to(acc);
    }
    if (bad) atomicOr(flags, FLAG_RANGE);
}

size_t dev_lincomb(Ctx* c, const Vec* const* vecs, const double* coef, int n, size_t cap, uint64_t* ok, double* ov) {
    size_t tot = 0;
    for (int i = 0; i < n; ++i) tot += (size_t)vecs[i]->len;
    if (tot == 0) return 0;
    uint64_t* ck = c->tmp_k[0].get<uint64_t>(tot);
    double* cv = c->tmp_v[0].get<double>(tot);
    uint64_t* ck2 = c->tmp_k[1].get<uint64_t>(tot);
    double* cv2 = c->tmp_v[1].get<double>(tot);
    size_t o = 0;
    for (int i = 0; i < n; ++i) {
        size_t l = (size_t)vecs[i]->len;
        if (l) k_scaled_copy<<<grid_of(l), SFCI_BLOCK, 0, c->stream>>>(vecs[i]->k, vecs[i]->v, l, coef[i], ck + o, cv + o);
        o += l;
    }
    uint64_t* kk = ck;
    double* vv = cv;
    radix_sort_pairs(c, &kk, &vv, ck2, cv2, tot, c->n_so);
    uint64_t* uk = (kk == ck) ? ck2 : ck;
    double* uv = (vv == cv) ? cv2 : cv;
    uint32_t* head = c->sel_b.get<uint32_t>(tot);
    uint64_t* pos = c->sel_c.get<uint64_t>(tot + 1);
    k_seg_heads<<<grid_of(tot), SFCI_BLOCK, 0, c->stream>>>(kk, tot, head);
    exclusive_scan_u32(c, head, tot, pos, pos + tot);
    k_seg_sum<<<grid_of(tot), SFCI_BLOCK, 0, c->stream>>>(kk, vv, tot, head, pos, uk, uv, c->d_flags);
    HIPCHK(hipGetLastError());
    uint64_t nu = 0;
    HIPCHK(hipMemcpyAsync(&nu, pos + tot, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (flags_read_clear(c) & FLAG_RANGE) throw HipError("lincomb: term outside the fixed-point range", -5);
    return select_top(c, uk, uv, nu, cap, ok, ov);
}

}  // namespace sfci
