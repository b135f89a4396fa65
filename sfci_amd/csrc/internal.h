// Internal declarations of the SFCI HIP library (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.h"
#include "../../include/sfci_gpu.h"

struct sfci_mlp;

namespace sfci {

struct HipError : std::runtime_error {
    int code;
    HipError(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess)                                                                      \
            throw ::sfci::HipError(std::string(#x) + ": " + hipGetErrorString(e_), -2);            \
    } while (0)

// Growable device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    template <typename T>
    T* get(size_t count) {
        size_t need = count * sizeof(T);
        if (need == 0) need = sizeof(T);
        if (need > bytes) {
            if (p) HIPCHK(hipFree(p));
            p = nullptr;
            bytes = 0;
            size_t alloc = need + need / 8;
            HIPCHK(hipMalloc(&p, alloc));
            bytes = alloc;
        }
        return (T*)p;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

// One candidate record: (pass-local dense target << 1 | is_diagonal, value), stored as
// one 16-byte (FP64) or 12-byte (FP32) structure so that a run of records is one
// contiguous span of memory.
struct alignas(16) Rec64 {
    uint64_t key;
    double val;
};
struct alignas(4) Rec32 {
    uint32_t klo, khi;
    float val;
};
template <typename T>
struct RecOf;
template <>
struct RecOf<double> {
    using type = Rec64;
};
template <>
struct RecOf<float> {
    using type = Rec32;
};
__host__ __device__ __forceinline__ uint64_t rec_key(const Rec64& r) { return r.key; }
__host__ __device__ __forceinline__ uint64_t rec_key(const Rec32& r) { return (uint64_t)r.khi << 32 | r.klo; }
__host__ __device__ __forceinline__ double rec_val(const Rec64& r) { return r.val; }
__host__ __device__ __forceinline__ float rec_val(const Rec32& r) { return r.val; }
__host__ __device__ __forceinline__ Rec64 make_rec(uint64_t k, double v) { return Rec64{k, v}; }
__host__ __device__ __forceinline__ Rec32 make_rec(uint64_t k, float v) {
    return Rec32{(uint32_t)k, (uint32_t)(k >> 32), v};
}
// Record lists are streamed: the generation kernel writes every FP64 record once, with a
// whole line per wavefront, through non-temporal stores, which keep the records from
// evicting the integral tables that the kernel rereads.  (Non-temporal loads, and the
// partition scatters' short runs, measured slower with the non-temporal forms, so they
// use ordinary accesses.)
typedef unsigned long long rec_u64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void rec_store(Rec64* p, const Rec64& r) {
    rec_u64x2 v;
    v.x = r.key;
    v.y = (unsigned long long)__double_as_longlong(r.val);
    __builtin_nontemporal_store(v, (rec_u64x2*)p);
}
__device__ __forceinline__ void rec_store(Rec32* p, const Rec32& r) { *p = r; }

// Compact records of a hash-merged target pass.  The pass-local dense target x (kbits bits)
// is replaced by its hash h = x * mul mod 2^kbits, a bijection (mul odd; inv is its inverse
// mod 2^kbits).  The top hb bits of h are the record's hash bucket, kept outside the record
// (a 16-bit side array at generation, implied by the record's position after the
// partition), and the record's 32-bit key is (h mod 2^rb) << 1 | is_diagonal, rb = kbits -
// hb <= 31.  A record is 12 bytes in FP64 and 8 in mixed precision.
struct CHash {
    int kbits, hb, rb;
    uint64_t mul, inv, mask;
};
__host__ __device__ __forceinline__ uint64_t ch_hash(const CHash& c, uint64_t x) { return (x * c.mul) & c.mask; }
__host__ __device__ __forceinline__ uint64_t ch_unhash(const CHash& c, uint64_t h) { return (h * c.inv) & c.mask; }
struct alignas(4) RecC64 {
    uint32_t key, vlo, vhi;
};
struct alignas(8) RecC32 {
    uint32_t key;
    float val;
};
template <typename T>
struct RecCOf;
template <>
struct RecCOf<double> {
    using type = RecC64;
};
template <>
struct RecCOf<float> {
    using type = RecC32;
};
__host__ __device__ __forceinline__ uint64_t rec_key(const RecC64& r) { return r.key; }
__host__ __device__ __forceinline__ uint64_t rec_key(const RecC32& r) { return r.key; }
__host__ __device__ __forceinline__ double rec_val(const RecC64& r) {
    const uint64_t b = (uint64_t)r.vhi << 32 | r.vlo;
    double v;
    __builtin_memcpy(&v, &b, 8);
    return v;
}
__host__ __device__ __forceinline__ float rec_val(const RecC32& r) { return r.val; }
__host__ __device__ __forceinline__ RecC64 make_crec(uint32_t k, double v) {
    uint64_t b;
    __builtin_memcpy(&b, &v, 8);
    return RecC64{k, (uint32_t)b, (uint32_t)(b >> 32)};
}
__host__ __device__ __forceinline__ RecC32 make_crec(uint32_t k, float v) { return RecC32{k, v}; }
// (three 4-byte stores: a 3-element vector type would be padded to 16 bytes)
__device__ __forceinline__ void rec_store(RecC64* p, const RecC64& r) {
    __builtin_nontemporal_store(r.key, &p->key);
    __builtin_nontemporal_store(r.vlo, &p->vlo);
    __builtin_nontemporal_store(r.vhi, &p->vhi);
}
__device__ __forceinline__ void rec_store(RecC32* p, const RecC32& r) { *p = r; }

// Candidate record of the 128-bit-key H.v: the target key (lo = spin orbitals 0-63, hi =
// 64-127) and its value; a diagonal record's value is a negative quiet NaN with bit 50 set
// whose low 50 bits are the source index (its FP64 diagonal term is kept per source).  No
// arithmetic result has that pattern (the default NaN is positive), and a product that is
// NaN or out of range is flagged where it is made.
struct alignas(8) RecW {
    uint64_t lo, hi;
    double val;
};
constexpr uint64_t WTAG_BITS = 0xfffc000000000000ull;
__host__ __device__ __forceinline__ double wtag_val(uint64_t j) {
    const uint64_t b = WTAG_BITS | j;
    double v;
    __builtin_memcpy(&v, &b, 8);
    return v;
}
// the source index of a diagonal record, or ~0 for an ordinary one
__host__ __device__ __forceinline__ uint64_t wtag_index(double v) {
    uint64_t b;
    __builtin_memcpy(&b, &v, 8);
    return (b & WTAG_BITS) == WTAG_BITS ? (b & 0x0003ffffffffffffull) : ~0ull;
}

// A diagonal (tag) record carries no value: its value field holds the bits of the source
// index j (< 2^32), so that the merge finds the source's FP64 diagonal term and input
// coefficient without a search.  Consumers that sum values treat tag records as zero.
template <typename T>
__host__ __device__ __forceinline__ T tag_val(uint32_t j);
template <>
__host__ __device__ __forceinline__ double tag_val<double>(uint32_t j) {
    const uint64_t b = j;
    double v;
    __builtin_memcpy(&v, &b, 8);
    return v;
}
template <>
__host__ __device__ __forceinline__ float tag_val<float>(uint32_t j) {
    float v;
    __builtin_memcpy(&v, &j, 4);
    return v;
}
__host__ __device__ __forceinline__ uint32_t tag_index(double v) {
    uint64_t b;
    __builtin_memcpy(&b, &v, 8);
    return (uint32_t)b;
}
__host__ __device__ __forceinline__ uint32_t tag_index(float v) {
    uint32_t b;
    __builtin_memcpy(&b, &v, 4);
    return b;
}

// Device-side error flags (bit set by kernels).
enum : uint32_t {
    FLAG_RANGE = 1u,        // a term left the fixed-point range (|x| >= 2^42)
    FLAG_OUT_OVERFLOW = 2u, // merged-output staging buffer too small (host grows and retries)
    FLAG_BAD_INPUT = 4u,    // an input determinant was not found in its own (unsorted) vector
    FLAG_BUCKET_FULL = 8u,  // a hash bucket had more distinct targets than its LDS table holds
    FLAG_REGION_FULL = 16u, // a first-level partition region received more records than it holds
    FLAG_GEN_OVERFLOW = 32u, // generation produced more records than the planning count
};

// Fixed per-Hamiltonian quantities passed by value to the kernels.
struct HamDev {
    int n_orb;
    int k;                 // electrons per spin (the two spins are equal)
    int v;                 // virtual spatial orbitals per spin
    uint64_t dim;          // C(n_orb, k): strings per spin
    const double* h;       // [2][n_orb][n_orb]  h[s][a][i] = h(2a+s, 2i+s)
    const double* g;       // [4][n_orb^4]       block 2*s1+s2: (a i | b j) with (a,i) spin s1, (b,j) spin s2
    uint64_t g_stride;     // elements between blocks: n_orb^4, or 0 when one block serves all four
    const double* gt;      // [2][n_orb^4] opposite-spin blocks as [a][i][j][b], b fastest: gt[0] holds
                           // (a i | b j) of block (0,1), gt[gt_stride] holds (b j | a i) of block (1,0)
    uint64_t gt_stride;
    const uint32_t* binom; // [33][33] C(n, t), n <= 32 (every entry is below 2^32)
    const uint16_t* occ_pairs;  // C(k,2) entries (lo byte, hi byte) = positions in the occupied list
    const uint16_t* vir_pairs;  // C(v,2) entries = positions in the virtual list
    double constant;
    // class sizes
    uint64_t n_sa, n_daa, n_dab, n_conn;   // singles per spin, same-spin doubles per spin, opposite-spin, U+1
    uint32_t n_opair, n_vpair;
};

// Same-spin double of one spin string (spin s, spatial i < j -> a < b).  Its element in a
// determinant whose other-spin string is T is (popc(T & m) odd ? -h : h): h is the element
// with the other spin empty (the magnitude and the sign from spin-s electrons depend only on
// the spin-s string), m the other-spin orbitals that lie strictly between an odd number of
// the excitation's two orbital pairs (the parity the other spin adds to both sign factors).
struct SsEnt {
    uint32_t r;   // colex rank of the excited spin-s string
    uint32_t m;
    double h;
};

// Same-spin double tables of one H.v, device pointers (a: spin 0, b: spin 1; null = the
// generator enumerates that class itself).  Spin 0: string u's entries at a + u * n_daa,
// grouped by target pass, pass p's at [a_start[u * (n_pass + 1) + p], ... + p + 1)
// (a_start null when n_pass == 1); a_idx[j] / b_idx[j]: source j's string index.
struct SsTabs {
    const SsEnt* a = nullptr;
    const uint32_t* a_start = nullptr;
    const uint32_t* a_idx = nullptr;
    const SsEnt* b = nullptr;
    const uint32_t* b_idx = nullptr;
    int pass = 0, n_pass = 1;
};

// Per-kernel timing with HIP events on the library stream (resolved after each call).
enum KernelId : int {
    K_GENERATE = 0,
    K_ROWGROUP,        // k_rowgroup_hist + k_count_passes: pass planning and per-source pass counts
    K_PART_HIST,       // k_part_hist*<DigitHash>: hash-partition histogram
    K_PART_SCATTER,    // k_part_scatter*<DigitHash>: second-level partition
    K_SORT_HIST,       // k_radix_hist<DigitPlain>: key sort histogram
    K_SORT_SCATTER,    // k_radix_scatter<DigitPlain>
    K_SCAN,
    K_BUCKET_BOUNDS,
    K_BUCKET_REDUCE,
    K_REDUCE,
    K_SELECT,
    K_VECOPS,
    K_ENERGY,
    K_PART_SCATTER1,   // k_part_scatter_cur<DigitHash>: first-level partition with region cursors
    K_MLP_SCORE,       // k_mlp_fwd on screened candidates (fused three-layer forward, BF16 MFMA)
    K_SCREEN_SELECT,   // score histograms and the kept-candidate compaction of the screened H.v
    K_MLP_TRAIN,       // one training step of the screening network (forward, backward, Adam)
    K_SS_TABLE,        // k_ss_table + string indexing: same-spin double tables of the source strings
    K_NUM
};

struct KernelStat {
    double ms = 0.0;
    uint64_t launches = 0;
    double bytes = 0.0;   // algorithmic bytes moved (reads + writes of the kernel's operands)
};

struct KernelClock {
    struct Rec {
        int id;
        double bytes;
        hipEvent_t a, b;
    };
    struct Phase {
        double* dst;
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    int open = 0;                  // phase timers alive (their events stay reserved)
    std::vector<Rec> recs;
    std::vector<Phase> phases;
    KernelStat stat[K_NUM];
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_so = 0, n_orb = 0, n_elec = 0;
    HamDev ham{};
    double* d_h = nullptr;
    double* d_g = nullptr;
    double* d_gt = nullptr;
    uint32_t* d_binom = nullptr;
    uint16_t* d_occ_pairs = nullptr;
    uint16_t* d_vir_pairs = nullptr;
    uint16_t* d_sdec = nullptr;      // single excitation x -> (occupied position | virtual position << 8)
    std::vector<uint64_t> h_binom;   // host copy
    double cut = 1e-5;
    size_t ws_limit = 0;             // bytes for the candidate buffers of one pass
    // scratch
    DevBuf cand_k[2], cand_v[2];
    DevBuf src_cnt, src_off, src_k, src_v, dterm;
    DevBuf stage_k, stage_v;         // merged H.v output before the top-cap select
    DevBuf lin_k[2], lin_v[2];       // uncut H.v results of the linear-solver products
    DevBuf part_meta[2], part_start[2], part_len[2], part_hbase[2], part_hstride[2], part_region[2];   // partition chunks
    DevBuf part_base;
    DevBuf part_dig;                 // second-level digit byte of each first-level partition slot
    DevBuf tmp_k[2], tmp_v[2];
    DevBuf sel_a, sel_b, sel_c;      // selection / merge scratch
    DevBuf scratch_hist, scratch_off, scratch_small;
    DevBuf scratch_scan[8];
    DevBuf rowhist;
    DevBuf plan_k[2], plan_v[2], plan_fl, plan_pos, plan_first, plan_str, plan_cnt;   // row-grouped planning
    DevBuf ss_tab[2], ss_start, ss_idx[2];   // same-spin double tables and per-source string indices
    DevBuf bstr_k[2], bstr_v[2], bstr_fl, bstr_pos, bstr_str;   // spin-1 string table
    int gen_tables = 0;              // same-spin double tables: 0 auto, 1 never, 2 always
    DevBuf bucket_bounds;
    double unique_ratio = 0.25;
    double pass_load = 0.9;          // distinct targets per LDS-table bucket aimed at, in table slots
    // distinct/candidate ratio of each target-row range (one per pass) of the previous H.v:
    // range lower bounds (ascending, the first 0) and ratios; the next call's planning weighs
    // each row segment by the ratio of its range (distinct targets are denser near the rows
    // of the largest entries)
    std::vector<uint64_t> rr_lo;
    std::vector<double> rr_ratio;
    bool g_one_block = false;        // the four two-electron spin blocks share one table
    int precision = 0;               // 0: FP64 candidate values; 1: FP32 values (mixed precision)
    int merge_mode = 0;              // 0 auto, 1 full sort only, 2 hash buckets with merge_hash_bits
    int merge_hash_bits = 0;         // hash bits of merge_mode 2
    int wide_merge_bits = 0;         // 128-bit-key H.v: hash bits that bucket records (0: from the pass size)
    int wide_fill = 0;               // 128-bit-key H.v: occupied slots at which a merge table splits (0: default)
    // candidate screening of sfci_hv (0 off, 1 by |value|, 2 by the network)
    int screen_mode = 0;
    const sfci_mlp* screen_net = nullptr;
    double screen_keep = 3.0;        // kept candidates per input determinant
    double screen_eps = 1e-7;        // prefilter |H_ij coeff c_j| > eps
    unsigned long long* d_acc = nullptr;   // 8 x 2 words of fx128 accumulators
    uint32_t* d_flags = nullptr;
    uint64_t* d_counter = nullptr;         // 8 general counters
    std::string last_error;
    KernelClock kclock;
    // statistics of the last H.v
    struct {
        uint64_t n_src = 0, n_candidates = 0, n_unique = 0, n_out = 0, n_passes = 0, n_fallback = 0;
        uint64_t n_split = 0;        // hash-bucket classes split because an LDS table filled
        uint64_t n_str0 = 0, n_str1 = 0;   // distinct source strings with same-spin tables (0: none)
        int hash_bits = 0;
        double ms_generate = 0, ms_sort = 0, ms_reduce = 0, ms_select = 0, ms_total = 0;
        uint64_t n_kept = 0;         // screened: candidates kept for the merge
        double screen_eps = 0.0, ms_score = 0.0;
    } st;
};

struct Vec {
    Ctx* ctx = nullptr;
    uint64_t* k = nullptr;
    double* v = nullptr;
    int64_t len = 0;
    int64_t cap = 0;
};

// sort.hip
void exclusive_scan_u32(Ctx* c, const uint32_t* in, size_t n, uint64_t* out, uint64_t* total_dev);
void exclusive_scan_u64(Ctx* c, const uint64_t* in, size_t n, uint64_t* out, uint64_t* total_dev);
void radix_sort_pairs(Ctx* c, uint64_t** keys, double** vals, uint64_t* ktmp, double* vtmp, size_t n, int bits);
template <typename T>
void radix_sort_kv(Ctx* c, uint64_t** keys, T** vals, uint64_t* ktmp, T* vtmp, size_t n, int bits);
template <typename T>
void radix_partition_hash(Ctx* c, uint64_t** keys, T** vals, uint64_t* ktmp, T* vtmp, size_t n, int hb);
// Group the n records of A by the hb-bit target hash (8 < hb <= 16: two passes, A -> B
// -> A; hb <= 8: one pass, A -> B), ascending in the hash, order inside a group
// unspecified.  B holds part_capacity(n) records.  Returns the grouped records, or
// nullptr when a pass-1 region overflowed (A is then unchanged).
size_t part_capacity(size_t n);
template <typename R>
R* hash_partition(Ctx* c, R* A, R* B, size_t n, int hb);
// The same grouping for compact records (8 < hb <= 16) whose hash buckets are in bkt:
// records end in A, dense in bucket order, with bucket b's records at [bstart[b], bend[b]).
// Returns false when a pass-1 region overflowed (A is then unchanged).
template <typename R>
bool hash_partition_c(Ctx* c, R* A, const uint16_t* bkt, R* B, size_t n, int hb, uint64_t* bstart, uint64_t* bend);

// noalloc.hip
// With `device` current: record that device's address of the selection scratch array.
void noalloc_setup(int device);
// select_top with the scratch array only (no device allocation); same result as select_top.
size_t dev_shrink_noalloc(Ctx* c, const uint64_t* k, const double* v, size_t n, size_t cap, uint64_t* ok,
                          double* ov);

// vecops.hip
// Keep the `cap` entries ranked highest by (|v| desc, key asc); order-preserving.
// Output may not alias input.  Returns the kept count.
size_t select_top(Ctx* c, const uint64_t* k, const double* v, size_t n, size_t cap, uint64_t* ok, double* ov);
double dev_dot(Ctx* c, const uint64_t* ak, const double* av, size_t na, const uint64_t* bk, const double* bv,
               size_t nb);
double dev_norm(Ctx* c, const double* v, size_t n);
void dev_scale(Ctx* c, double* v, size_t n, double s);
size_t dev_add_two(Ctx* c, const uint64_t* vk, const double* vv, size_t nv, const uint64_t* wk, const double* wv,
                   size_t nw, double a, size_t cap, uint64_t* ok, double* ov);
size_t dev_lincomb(Ctx* c, const Vec* const* vecs, const double* coef, int n, size_t cap, uint64_t* ok, double* ov);
void dev_copy(Ctx* c, uint64_t* dk, double* dv, const uint64_t* sk, const double* sv, size_t n);
// Order-preserving compaction of the entries with |v| > cut (strict).  Returns the count.
size_t dev_filter_above(Ctx* c, const uint64_t* k, const double* v, size_t n, double cut, uint64_t* ok, double* ov);

// hv.hip
// H.v with the reference's semantics; returns the output length written to (ok, ov)
// (capacity out_cap, key-sorted) and the exact diag.
size_t dev_hv(Ctx* c, const uint64_t* ik, const double* iv, size_t n_in, size_t nd, double lambda, double coeff,
              size_t out_cap, uint64_t* ok, double* ov, double* diag);
// Every distinct target of coeff (H - lambda) v, no cut and no cap (exact zeros kept),
// key-sorted.  The result lives in library buffers that the next H.v call reuses.
struct HvAll {
    const uint64_t* k = nullptr;
    const double* v = nullptr;
    size_t n = 0;
};
void dev_hv_all(Ctx* c, const uint64_t* ik, const double* iv, size_t n_in, double lambda, double coeff, double* diag,
                HvAll* all);
// H.v on 128-bit keys (wide.hip): keys are two words per entry (lo = spin orbitals 0-63,
// hi = 64-127), ascending as hi:lo; same semantics and outputs as dev_hv.
size_t dev_hv_wide(Ctx* c, int n_so, const uint64_t* ik2, const double* iv, size_t n_in, size_t nd, double lambda,
                   double coeff, size_t out_cap, uint64_t* ok2, double* ov, double* diag);
// Vector operations on two-word keys (wide.hip), same semantics as their 64-bit forms.
double dev_wide_dot(Ctx* c, const uint64_t* ak2, const double* av, size_t na, const uint64_t* bk2, const double* bv,
                    size_t nb);
size_t dev_wide_add_two(Ctx* c, const uint64_t* vk2, const double* vv, size_t nv, const uint64_t* wk2,
                        const double* wv, size_t nw, double a, size_t cap, uint64_t* ok2, double* ov);
size_t dev_wide_shrink(Ctx* c, const uint64_t* sk2, const double* sv, size_t n, size_t cap, uint64_t* ok2,
                       double* ov);
double dev_wide_energy(Ctx* c, const uint64_t* k2, const double* v, size_t n);
// <v|H|v> with calc_energy's per-source pull (no candidate list); v key-sorted.
double dev_energy(Ctx* c, const uint64_t* k, const double* v, size_t n);
// cisd.hip: the CISD block eigenvector (state_idx, ascending) of ref_det into (ok, ov),
// key order; returns the block size.
size_t dev_cisd(Ctx* c, uint64_t ref_det, int state_idx, size_t out_cap, uint64_t* ok, double* ov,
                double* eigenvalue);
// Every connection of each of the ns sources with unit amplitude: record j * n_conn + e
// holds ((dense target) << 1 | is_diagonal, H_target,source); the diagonal records carry
// value 0 and dterm[j] = H_jj.  Buffers are the library's candidate buffers.
void dev_connection_records(Ctx* c, const uint64_t* sk, size_t ns, const Rec64** rec, const double** dterm);

// ---- candidate screening (screen.hip / mlp.hip) ------------------------------------
// A screened candidate: target (dense index << 1 | is_diagonal, as in Rec64), the target
// determinant's bits, its FP32 value H_ij * coeff * c_j (the source index bits for a
// diagonal tag), and the excitation rank of the connection (0 diagonal, 1 single, 2 double).
struct alignas(8) RecS {
    uint64_t key;
    uint64_t det;
    float val;
    uint32_t cls;
};
#if defined(__HIPCC__)
// Stores of a RecS through pointers of a known address space (LDS: 3, global: 1): through
// a generic pointer the compiler emits flat stores.
template <int AS>
__device__ __forceinline__ void recs_store(RecS* p, const RecS& r) {
    __attribute__((address_space(AS))) RecS* q = (__attribute__((address_space(AS))) RecS*)p;
    q->key = r.key;
    q->det = r.det;
    q->val = r.val;
    q->cls = r.cls;
}
#endif

// Screening feature of a candidate value: log2|v| piecewise linear in the mantissa,
// exponent + mantissa fraction of the float (one IEEE rounding; -150 for zero).
__host__ __device__ __forceinline__ float plog2(float v) {
    uint32_t b;
    __builtin_memcpy(&b, &v, 4);
    b &= 0x7fffffffu;
    if (b == 0) return -150.0f;
    return (float)((int)(b >> 23) - 127) + (float)(b & 0x7fffffu) * (1.0f / 8388608.0f);
}
// Screened H.v (Ctx::screen_mode != 0) with the semantics of sfci_hv on the kept candidates.
size_t dev_hv_screened(Ctx* c, const uint64_t* ik, const double* iv, size_t n_in, size_t nd, double lambda,
                       double coeff, size_t out_cap, uint64_t* ok, double* ov, double* diag);
// Training samples from the candidates of v (see sfci_screen_samples); returns the count.
size_t dev_screen_samples(Ctx* c, const uint64_t* vk, const double* vv, size_t n, double eps, size_t max_samples,
                          uint64_t seed, sfci_screen_sample_t* out);

// Screening network 66 -> 256 -> 256 -> 1 (SiLU), BF16 operands, FP32 accumulation.
constexpr int MLP_IN = 66, MLP_INP = 80, MLP_HID = 256;
constexpr int MLP_OFF_W1 = 0, MLP_OFF_B1 = MLP_HID * MLP_IN, MLP_OFF_W2 = MLP_OFF_B1 + MLP_HID,
              MLP_OFF_B2 = MLP_OFF_W2 + MLP_HID * MLP_HID, MLP_OFF_W3 = MLP_OFF_B2 + MLP_HID,
              MLP_OFF_B3 = MLP_OFF_W3 + MLP_HID, MLP_NPARAM = MLP_OFF_B3 + 1;
// score[i] = network output for record i's features (with_magnitude: the network's predicted
// ln|y| of the target plus ln|v| of the record); diagonal tags get +inf.
void mlp_score_records(Ctx* c, const sfci_mlp* net, const RecS* rec, size_t n, float* score, bool with_magnitude);

// kernel timing: kt_begin before a launch, kt_end after it; kt_resolve synchronizes the
// stream and folds the elapsed times into c->kclock.stat (and the phase timers' laps into
// their destinations).
int kt_begin(Ctx* c, int id, double bytes);
void kt_end(Ctx* c, int rec);
void kt_resolve(Ctx* c);

// Phase timer on the library stream.  lap(dst) records the end of the phase that began at
// construction or at the previous lap, and starts the next one; the elapsed time is added
// to *dst when the clock is resolved (kt_resolve), so timing a phase adds no host
// synchronization.  dst must stay valid until then.
struct PhaseTimer {
    Ctx* c;
    hipEvent_t a;
    explicit PhaseTimer(Ctx* ctx);
    ~PhaseTimer();
    void lap(double* dst);
    PhaseTimer(const PhaseTimer&) = delete;
    PhaseTimer& operator=(const PhaseTimer&) = delete;
};

// accumulator helpers
void acc_clear(Ctx* c, int slot);
fx128 acc_read_raw(Ctx* c, int slot);
double acc_read(Ctx* c, int slot);
uint32_t flags_read_clear(Ctx* c);
void atomic_flags_restore(Ctx* c, uint32_t f);   // put back flags read by flags_read_clear

// Exact accumulation without a carry chain: a term x (|x| < 2^42) is split into three
// signed limbs with x * 2^84 = l2 * 2^84 + l1 * 2^42 + l0 truncated toward zero (the same
// integer fx_from gives), using exact FP64 steps: l2 = trunc(x), l1 = trunc(frac * 2^42),
// l0 = trunc(rest * 2^42).  Every limb has the sign of x and |l1|, |l0| < 2^42.  Each limb
// is added to its own 64-bit LDS word with an atomic whose result is unused.  A target
// receives fewer than 2^21 terms (at most one per source), so no limb sum overflows; the
// limbs are recombined exactly when the target is finished.  An integer t with |t| < 2^51
// converts to int64 as the bits of t + 1.5 * 2^52 minus those of 1.5 * 2^52 (the sum is exact
// and its ulp is 1), one add and one integer subtract instead of the generic conversion; the
// integer part is formed only when |x| >= 1.
__device__ __forceinline__ void limb_split(double x, unsigned long long& l0, unsigned long long& l1,
                                           unsigned long long& l2) {
    constexpr double MAGIC = 0x1.8p52;
    const long long mb = __double_as_longlong(MAGIC);
    double r1;
    l2 = 0;
    if (fabs(x) < 1.0) {
        r1 = x * 0x1p42;
    } else {
        if (!(fabs(x) < 4398046511104.0)) x = 0.0;   // out of range: flagged where the term was made
        const double t2 = trunc(x);
        l2 = (unsigned long long)(__double_as_longlong(t2 + MAGIC) - mb);
        r1 = (x - t2) * 0x1p42;
    }
    const double t1 = trunc(r1);
    const double t0 = trunc((r1 - t1) * 0x1p42);
    l1 = (unsigned long long)(__double_as_longlong(t1 + MAGIC) - mb);
    l0 = (unsigned long long)(__double_as_longlong(t0 + MAGIC) - mb);
}

__device__ __forceinline__ fx128 limb_join(unsigned long long l0, unsigned long long l1, unsigned long long l2) {
    fx128 t{l0, (long long)l0 < 0 ? ~0ull : 0ull};
    t = fx_add(t, fx128{l1 << 42, (unsigned long long)((long long)l1 >> 22)});
    return fx_add(t, fx128{0, l2 << 20});
}

// Subclasses for a class whose table filled: its targets D estimated from the fill after
// n_read of its records (records drawn over D keys: fill = D (1 - exp(-n_read / D)), solved
// by bisection), at `target` targets per subclass, at most `stack`.  Kept out of line: it
// runs once per split.
static __device__ __attribute__((noinline)) uint32_t split_count_for(uint32_t fill_u, uint32_t nread_u, int target,
                                                                     int stack) {
    const double fill = (double)fill_u, nread = (double)nread_u;
    double lo = fill, hi = fmax(fill, 64.0 * nread);
    double D = hi;
    if (nread > fill + 0.5) {
        for (int it = 0; it < 60; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (mid * (1.0 - exp(-nread / mid)) < fill) lo = mid;
            else hi = mid;
        }
        D = hi;
    }
    return (uint32_t)fmin(ceil(D / (double)target), (double)stack);
}

static inline int bits_for(uint64_t max_value_exclusive) {
    // smallest b with 2^b >= max_value_exclusive
    int b = 0;
    while (b < 64 && ((uint64_t)1 << b) < max_value_exclusive) ++b;
    return b;
}

}  // namespace sfci

// Handles of the C-ABI (opaque to callers).
struct sfci_ctx {
    sfci::Ctx c;
};
struct sfci_vec {
    sfci::Vec v;
};
// A context for more than 64 spin orbitals (128-bit keys) and its vectors: keys hold two
// words per entry.
struct sfci_wctx {
    sfci_ctx* x;
};
struct sfci_wvec {
    sfci::Ctx* ctx = nullptr;
    uint64_t* k = nullptr;   // 2 * cap words
    double* v = nullptr;
    int64_t len = 0;
    int64_t cap = 0;
};
namespace sfci {
size_t dev_wide_lincomb(Ctx* c, const sfci_wvec* const* vecs, const double* coef, int nvec, size_t cap,
                        uint64_t* ok2, double* ov);
}  // namespace sfci

namespace sfci {

inline void set_err(Ctx* c, const std::string& m) {
    if (c) c->last_error = m;
}

// Runs f on the context's device and turns exceptions into C-ABI return codes
// (HipError's code, SFCI_E_HIP (-2) for a failed host allocation, SFCI_E_ARG (-1) otherwise).
template <class F>
int guarded(Ctx* c, F&& f) {
    try {
        if (c) HIPCHK(hipSetDevice(c->device));
        f();
        return 0;
    } catch (const HipError& e) {
        set_err(c, e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_err(c, "host allocation failed");
        return -2;
    } catch (const std::exception& e) {
        set_err(c, e.what());
        return -1;
    }
}

}  // namespace sfci
