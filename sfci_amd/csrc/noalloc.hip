// Top-cap selection that allocates no device memory (sfci_shrink_from_noalloc).
//
// The same result as sfci_shrink_from (shrinkFrom, src/all.h:281-299, with the declared tie
// rule: larger |c| first, equal |c| by smaller key), for callers that cannot allocate: code
// under stream capture, tight serving loops.  Everything the selection needs beyond the two
// vectors lives in `scratch`, a file-scope device array of NA_WORDS 64-bit words (18.5 KiB)
// that every GPU holds for the life of the process:
//   [0, 256)                 digit histogram of the current radix-select round
//   [256, 256 + NA_CHUNKS)   kept entries per input chunk (the compaction's block offsets)
// Host code never passes `scratch` itself to a kernel: its address on each device comes from
// hipGetSymbolAddress, taken once per device when a context is created (noalloc_setup).
// Every call on one GPU shares that GPU's array, so calls that overlap on one GPU race.
//
// Method: an MSD radix select on the |c| bit pattern (8-bit digits) finds tau, the |c| of the
// cap-th entry, and how many entries equal to tau are kept; a second radix select on the keys
// of those tied entries finds the largest kept key; then an order-preserving compaction in
// NA_CHUNKS contiguous chunks (a count pass, then a scatter pass whose chunk offsets are sums
// of the counts) writes the kept entries in key order.
#include <mutex>

#include "internal.h"

namespace sfci {

constexpr int NA_CHUNKS = 2048;
constexpr int NA_HIST = 0;
constexpr int NA_COUNT = 256;
constexpr int NA_WORDS = NA_COUNT + NA_CHUNKS;
constexpr int NA_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned long long scratch[NA_WORDS];

static unsigned long long* g_scratch_ptr[NA_MAX_DEVICES];
static std::mutex g_scratch_mu;

// With `device` current: record this device's address of `scratch` (once per device).
void noalloc_setup(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) throw HipError("noalloc: device index out of range", SFCI_E_ARG);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw HipError(std::string("noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e), SFCI_E_HIP);
    g_scratch_ptr[device] = (unsigned long long*)p;
}

static unsigned long long* scratch_of(int device) {
    unsigned long long* p = nullptr;
    if (device >= 0 && device < NA_MAX_DEVICES) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        p = g_scratch_ptr[device];
    }
    if (!p) throw HipError("noalloc: the scratch address of this device was not set up", SFCI_E_HIP);
    return p;
}

__device__ __forceinline__ uint64_t na_abs_bits(double x) {
    return (uint64_t)__double_as_longlong(x) & 0x7fffffffffffffffull;
}

__global__ void k_na_zero(unsigned long long* __restrict__ p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

__global__ void k_na_copy(const uint64_t* __restrict__ k, const double* __restrict__ v, size_t n,
                          uint64_t* __restrict__ ok, double* __restrict__ ov) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        ok[i] = k[i];
        ov[i] = v[i];
    }
}

// digit (|v| >> shift) & 255 of the entries whose |v| bits match prefix under mask
__global__ __launch_bounds__(SFCI_BLOCK) void k_na_hist_abs(const double* __restrict__ v, size_t n, uint64_t prefix,
                                                            uint64_t mask, int shift,
                                                            unsigned long long* __restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * SFCI_BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * SFCI_BLOCK) {
        const uint64_t b = na_abs_bits(v[i]);
        if ((b & mask) == prefix) atomicAdd(&h[(b >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// digit (key >> shift) & 255 of the entries with |v| bits tau whose key matches kprefix under kmask
__global__ __launch_bounds__(SFCI_BLOCK) void k_na_hist_key(const uint64_t* __restrict__ k,
                                                            const double* __restrict__ v, size_t n, uint64_t tau,
                                                            uint64_t kprefix, uint64_t kmask, int shift,
                                                            unsigned long long* __restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * SFCI_BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * SFCI_BLOCK) {
        if (na_abs_bits(v[i]) != tau) continue;
        const uint64_t key = k[i];
        if ((key & kmask) == kprefix) atomicAdd(&h[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

__device__ __forceinline__ bool na_kept(uint64_t key, double x, uint64_t tau, uint64_t key_cut) {
    const uint64_t b = na_abs_bits(x);
    return b > tau || (b == tau && key <= key_cut);
}

// one workgroup per chunk [b * chunk, min(n, (b + 1) * chunk)): its kept count
__global__ __launch_bounds__(SFCI_BLOCK) void k_na_count(const uint64_t* __restrict__ k, const double* __restrict__ v,
                                                         size_t n, size_t chunk, uint64_t tau, uint64_t key_cut,
                                                         unsigned long long* __restrict__ counts) {
    __shared__ uint32_t s_w[SFCI_BLOCK / SFCI_WAVE];
    const size_t lo = (size_t)blockIdx.x * chunk, hi = min(n, lo + chunk);
    uint32_t c = 0;
    for (size_t i = lo + threadIdx.x; i < hi; i += SFCI_BLOCK) c += na_kept(k[i], v[i], tau, key_cut) ? 1u : 0u;
    for (int o = SFCI_WAVE / 2; o > 0; o >>= 1) c += __shfl_down(c, o, SFCI_WAVE);
    if ((threadIdx.x & (SFCI_WAVE - 1)) == 0) s_w[threadIdx.x / SFCI_WAVE] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < SFCI_BLOCK / SFCI_WAVE; ++w) t += s_w[w];
        counts[blockIdx.x] = t;
    }
}

// one workgroup per chunk: kept entries at (sum of the earlier chunks' counts) + their rank
// inside the chunk, tile by tile in input order
__global__ __launch_bounds__(SFCI_BLOCK) void k_na_scatter(const uint64_t* __restrict__ k,
                                                           const double* __restrict__ v, size_t n, size_t chunk,
                                                           uint64_t tau, uint64_t key_cut,
                                                           const unsigned long long* __restrict__ counts,
                                                           uint64_t* __restrict__ ok, double* __restrict__ ov) {
    __shared__ unsigned long long s_part[SFCI_BLOCK / SFCI_WAVE];
    __shared__ uint32_t s_w[SFCI_BLOCK / SFCI_WAVE];
    const int tid = threadIdx.x, lane = tid & (SFCI_WAVE - 1), w = tid / SFCI_WAVE;
    unsigned long long pre = 0;
    for (uint32_t b = tid; b < blockIdx.x; b += SFCI_BLOCK) pre += counts[b];
    for (int o = SFCI_WAVE / 2; o > 0; o >>= 1) pre += __shfl_down(pre, o, SFCI_WAVE);
    if (lane == 0) s_part[w] = pre;
    __syncthreads();
    unsigned long long base = 0;
    for (int i = 0; i < SFCI_BLOCK / SFCI_WAVE; ++i) base += s_part[i];
    const size_t lo = (size_t)blockIdx.x * chunk, hi = min(n, lo + chunk);
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (size_t t0 = lo; t0 < hi; t0 += SFCI_BLOCK) {
        const size_t i = t0 + tid;
        bool keep = false;
        uint64_t key = 0;
        double x = 0.0;
        if (i < hi) {
            key = k[i];
            x = v[i];
            keep = na_kept(key, x, tau, key_cut);
        }
        const uint64_t bal = __ballot(keep);
        if (lane == 0) s_w[w] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int q = 0; q < SFCI_BLOCK / SFCI_WAVE; ++q) {
            if (q < w) before += s_w[q];
            total += s_w[q];
        }
        if (keep) {
            const uint64_t at = base + before + (uint32_t)__popcll(bal & lt);
            ok[at] = key;
            ov[at] = x;
        }
        base += total;
        __syncthreads();
    }
}

static void na_hist_read(Ctx* c, unsigned long long* hist, unsigned long long (&h)[256]) {
    HIPCHK(hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
}

size_t dev_shrink_noalloc(Ctx* c, const uint64_t* k, const double* v, size_t n, size_t cap, uint64_t* ok,
                          double* ov) {
    if (n > SFCI_NOALLOC_MAX_N)
        throw HipError("shrink_from_noalloc: source longer than SFCI_NOALLOC_MAX_N entries", SFCI_E_CAPACITY);
    unsigned long long* scr = scratch_of(c->device);
    const int grid = (int)std::max<size_t>(1, std::min<size_t>((n + SFCI_BLOCK - 1) / SFCI_BLOCK, 1024));
    if (n <= cap) {
        if (n) k_na_copy<<<grid, SFCI_BLOCK, 0, c->stream>>>(k, v, n, ok, ov);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        return n;
    }
    if (cap == 0) return 0;
    unsigned long long* hist = scr + NA_HIST;
    unsigned long long h[256];
    // tau: the |c| bit pattern of the cap-th largest entry; need: how many entries equal to
    // tau are kept
    uint64_t prefix = 0, mask = 0, need = cap;
    for (int shift = 56; shift >= 0; shift -= 8) {
        k_na_zero<<<1, 256, 0, c->stream>>>(hist, 256);
        k_na_hist_abs<<<grid, SFCI_BLOCK, 0, c->stream>>>(v, n, prefix, mask, shift, hist);
        HIPCHK(hipGetLastError());
        na_hist_read(c, hist, h);
        uint64_t above = 0;
        int d = 255;
        for (; d >= 0; --d) {
            if (above + h[d] >= need) break;
            above += h[d];
        }
        if (d < 0) throw HipError("shrink_from_noalloc: inconsistent histogram", SFCI_E_ARG);
        need -= above;
        prefix |= (uint64_t)d << shift;
        mask |= (uint64_t)255 << shift;
    }
    const uint64_t tau = prefix;
    // key_cut: the need-th smallest key among the entries equal to tau (all of them when
    // need equals their number)
    uint64_t kprefix = 0, kmask = 0, kneed = need, n_ties = 0;
    for (int shift = 56; shift >= 0; shift -= 8) {
        k_na_zero<<<1, 256, 0, c->stream>>>(hist, 256);
        k_na_hist_key<<<grid, SFCI_BLOCK, 0, c->stream>>>(k, v, n, tau, kprefix, kmask, shift, hist);
        HIPCHK(hipGetLastError());
        na_hist_read(c, hist, h);
        if (shift == 56) {
            for (int d = 0; d < 256; ++d) n_ties += h[d];
            if (n_ties == need) break;
        }
        uint64_t below = 0;
        int d = 0;
        for (; d < 256; ++d) {
            if (below + h[d] >= kneed) break;
            below += h[d];
        }
        if (d == 256) throw HipError("shrink_from_noalloc: inconsistent key histogram", SFCI_E_ARG);
        kneed -= below;
        kprefix |= (uint64_t)d << shift;
        kmask |= (uint64_t)255 << shift;
    }
    const uint64_t key_cut = (n_ties == need) ? ~0ull : kprefix;
    const size_t chunk = (n + NA_CHUNKS - 1) / NA_CHUNKS;
    const unsigned n_chunks = (unsigned)((n + chunk - 1) / chunk);
    unsigned long long* counts = scr + NA_COUNT;
    k_na_count<<<n_chunks, SFCI_BLOCK, 0, c->stream>>>(k, v, n, chunk, tau, key_cut, counts);
    k_na_scatter<<<n_chunks, SFCI_BLOCK, 0, c->stream>>>(k, v, n, chunk, tau, key_cut, counts, ok, ov);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return cap;
}

}  // namespace sfci
