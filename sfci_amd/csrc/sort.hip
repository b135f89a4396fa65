// Device-wide exclusive scan and stable LSD radix sort of (u64 key, f64 value) pairs.
//
// Radix sort: 8-bit digits, tiles of 4096 items (256 threads x 16 rounds).  Per pass:
//   hist    - per-tile digit histogram in LDS, written digit-major [256][n_tiles]
//   scan    - exclusive scan of the histogram -> global start of (digit, tile)
//   scatter - stable in-tile ranking: each round of 256 consecutive items is ranked
//             with a 64-lane digit match built from 8 ballots, wave counts are
//             combined in wave order through LDS, so items keep their input order.
// Passes run back to back on the stream with no host synchronization.
#include "internal.h"

namespace sfci {

constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SFCI_BLOCK * SCAN_ITEMS;   // 1024

// block-wide exclusive scan of one value per thread; returns the block total
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T& total) {
    __shared__ T s_w[SFCI_BLOCK / SFCI_WAVE];
    int lane = threadIdx.x & (SFCI_WAVE - 1), w = threadIdx.x / SFCI_WAVE;
    T x = v;
    for (int off = 1; off < SFCI_WAVE; off <<= 1) {
        T y = __shfl_up(x, off, SFCI_WAVE);
        if (lane >= off) x += y;
    }
    if (lane == SFCI_WAVE - 1) s_w[w] = x;
    __syncthreads();
    T pre = 0, tot = 0;
    for (int i = 0; i < SFCI_BLOCK / SFCI_WAVE; ++i) {
        if (i < w) pre += s_w[i];
        tot += s_w[i];
    }
    __syncthreads();
    total = tot;
    return pre + x - v;
}

template <typename TIn>
__global__ __launch_bounds__(SFCI_BLOCK) void k_scan_reduce(const TIn* in, size_t n, uint64_t* partial) {
    size_t base = (size_t)blockIdx.x * SCAN_TILE;
    uint64_t s = 0;
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        size_t i = base + (size_t)threadIdx.x * SCAN_ITEMS + r;
        if (i < n) s += (uint64_t)in[i];
    }
    uint64_t tot;
    block_exclusive_scan<uint64_t>(s, tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

template <typename TIn>
__global__ __launch_bounds__(SFCI_BLOCK) void k_scan_down(const TIn* in, size_t n, const uint64_t* partial_scanned,
                                                          uint64_t* out, uint64_t* total_out) {
    size_t base = (size_t)blockIdx.x * SCAN_TILE;
    uint64_t v[SCAN_ITEMS];
    uint64_t s = 0;
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        size_t i = base + (size_t)threadIdx.x * SCAN_ITEMS + r;
        v[r] = i < n ? (uint64_t)in[i] : 0;
        s += v[r];
    }
    uint64_t tot;
    uint64_t pre = block_exclusive_scan<uint64_t>(s, tot) + partial_scanned[blockIdx.x];
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        size_t i = base + (size_t)threadIdx.x * SCAN_ITEMS + r;
        if (i < n) out[i] = pre;
        pre += v[r];
    }
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == SFCI_BLOCK - 1) *total_out = pre;
}

template <typename TIn>
static void scan_impl(Ctx* c, const TIn* in, size_t n, uint64_t* out, uint64_t* total_dev, int depth) {
    if (n == 0) {
        if (total_dev) HIPCHK(hipMemsetAsync(total_dev, 0, 8, c->stream));
        return;
    }
    size_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    uint64_t* partial = c->scratch_scan[depth].get<uint64_t>(2 * nb + 2);
    uint64_t* pscan = partial + nb + 1;
    k_scan_reduce<TIn><<<nb, SFCI_BLOCK, 0, c->stream>>>(in, n, partial);
    if (nb == 1) {
        HIPCHK(hipMemsetAsync(pscan, 0, 8, c->stream));
    } else {
        scan_impl<uint64_t>(c, partial, nb, pscan, nullptr, depth + 1);
    }
    k_scan_down<TIn><<<nb, SFCI_BLOCK, 0, c->stream>>>(in, n, pscan, out, total_dev);
}

void exclusive_scan_u32(Ctx* c, const uint32_t* in, size_t n, uint64_t* out, uint64_t* total_dev) {
    scan_impl<uint32_t>(c, in, n, out, total_dev, 0);
}
void exclusive_scan_u64(Ctx* c, const uint64_t* in, size_t n, uint64_t* out, uint64_t* total_dev) {
    scan_impl<uint64_t>(c, in, n, out, total_dev, 0);
}

// ---------------------------------------------------------------------------------
constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;
constexpr int SORT_ROUNDS = 16;
constexpr int SORT_TILE = SFCI_BLOCK * SORT_ROUNDS;   // 4096

// digit of a key: plain bits, or bits of the target hash (key >> 1 hashed, hb bits)
struct DigitPlain {
    static constexpr int kHist = K_SORT_HIST, kScatter = K_SORT_SCATTER;
    int shift;
    __device__ __forceinline__ uint32_t operator()(uint64_t k) const { return (uint32_t)(k >> shift) & (RADIX - 1); }
};
struct DigitHash {
    static constexpr int kHist = K_PART_HIST, kScatter = K_PART_SCATTER;
    int shift, hb;
    __device__ __forceinline__ uint32_t operator()(uint64_t k) const {
        return (uint32_t)(target_hash(k >> 1, hb) >> shift) & (RADIX - 1);
    }
};

template <class Digit>
__global__ __launch_bounds__(SFCI_BLOCK) void k_radix_hist(const uint64_t* __restrict__ kin, size_t n, Digit dg,
                                                           uint32_t* __restrict__ hist, size_t n_tiles) {
    __shared__ uint32_t h[RADIX];
    h[threadIdx.x] = 0;
    __syncthreads();
    size_t base = (size_t)blockIdx.x * SORT_TILE;
    for (int r = 0; r < SORT_ROUNDS; ++r) {
        size_t i = base + (size_t)r * SFCI_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&h[dg(kin[i])], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile (SORT_TILE items, SCATTER_THREADS threads).  Items are
// ranked round by round (one item per thread per round, 8 ballots build each lane's
// digit-peer mask, wave counts are combined in wave order), placed in LDS in tile-local
// digit order, then written out so that consecutive threads write consecutive slots of
// each digit's output run.
constexpr int SCATTER_THREADS = 512;
constexpr int SCATTER_WAVES = SCATTER_THREADS / SFCI_WAVE;
constexpr int SCATTER_ROUNDS = SORT_TILE / SCATTER_THREADS;

template <class Digit, typename T>
__global__ __launch_bounds__(SCATTER_THREADS) void k_radix_scatter(const uint64_t* __restrict__ kin,
                                                                   const T* __restrict__ vin,
                                                                   uint64_t* __restrict__ kout,
                                                                   T* __restrict__ vout, size_t n, Digit dg,
                                                                   const uint64_t* __restrict__ off,
                                                                   const uint32_t* __restrict__ hist, size_t n_tiles) {
    __shared__ uint64_t s_key[SORT_TILE];
    __shared__ T s_val[SORT_TILE];
    __shared__ uint64_t s_gbase[RADIX];     // global start of each digit's run for this tile
    __shared__ uint32_t s_lstart[RADIX];    // tile-local start of each digit
    __shared__ uint32_t s_run[RADIX];       // items of each digit placed so far
    __shared__ uint32_t s_cnt[SCATTER_WAVES][RADIX];
    __shared__ uint32_t s_wsum[RADIX / SFCI_WAVE];
    const int tid = threadIdx.x, lane = tid & (SFCI_WAVE - 1), w = tid / SFCI_WAVE;
    const size_t tile = blockIdx.x;
    uint32_t h = 0, x = 0;
    if (tid < RADIX) {
        // exclusive scan of the tile histogram over digits (one digit per thread)
        h = hist[(size_t)tid * n_tiles + tile];
        x = h;
        for (int o = 1; o < SFCI_WAVE; o <<= 1) {
            uint32_t y = __shfl_up(x, o, SFCI_WAVE);
            if (lane >= o) x += y;
        }
        if (lane == SFCI_WAVE - 1) s_wsum[w] = x;
    }
    for (int i = tid; i < SCATTER_WAVES * RADIX; i += SCATTER_THREADS) (&s_cnt[0][0])[i] = 0;
    __syncthreads();
    if (tid < RADIX) {
        uint32_t pre = 0;
        for (int j = 0; j < w; ++j) pre += s_wsum[j];
        s_lstart[tid] = pre + x - h;
        s_gbase[tid] = off[(size_t)tid * n_tiles + tile];
        s_run[tid] = 0;
    }
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t base = tile * SORT_TILE;
    const int n_here = (int)((n - base) < (size_t)SORT_TILE ? (n - base) : (size_t)SORT_TILE);
    for (int r = 0; r < SCATTER_ROUNDS; ++r) {
        int li = r * SCATTER_THREADS + tid;
        bool valid = li < n_here;
        uint64_t k = valid ? kin[base + li] : 0ull;
        T v = valid ? vin[base + li] : (T)0;
        uint32_t d = dg(k);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < RADIX_BITS; ++b) {
            uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        uint32_t rank = (uint32_t)popc64(peers & lt);
        if (valid && rank == 0) s_cnt[w][d] = (uint32_t)popc64(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = s_lstart[d] + s_run[d] + rank;
            for (int j = 0; j < w; ++j) pos += s_cnt[j][d];
            s_key[pos] = k;
            s_val[pos] = v;
        }
        __syncthreads();
        if (tid < RADIX) {
            uint32_t tot = 0;
            for (int j = 0; j < SCATTER_WAVES; ++j) {
                tot += s_cnt[j][tid];
                s_cnt[j][tid] = 0;
            }
            s_run[tid] += tot;
        }
        __syncthreads();
    }
    for (int p = tid; p < n_here; p += SCATTER_THREADS) {
        uint64_t k = s_key[p];
        uint32_t d = dg(k);
        uint64_t g = s_gbase[d] + (uint64_t)(p - s_lstart[d]);
        kout[g] = k;
        vout[g] = s_val[p];
    }
}

template <typename T, class MakeDigit>
static void lsd_passes(Ctx* c, uint64_t** keys, T** vals, uint64_t* ktmp, T* vtmp, size_t n, int bits,
                       MakeDigit make) {
    if (n <= 1 || bits <= 0) return;
    size_t n_tiles = (n + SORT_TILE - 1) / SORT_TILE;
    uint32_t* hist = c->scratch_hist.get<uint32_t>((size_t)RADIX * n_tiles);
    uint64_t* off = c->scratch_off.get<uint64_t>((size_t)RADIX * n_tiles + 1);
    uint64_t* kin = *keys;
    T* vin = *vals;
    uint64_t* kout = ktmp;
    T* vout = vtmp;
    for (int shift = 0; shift < bits; shift += RADIX_BITS) {
        auto dg = make(shift);
        int t0 = kt_begin(c, decltype(dg)::kHist, 8.0 * n + 4.0 * RADIX * n_tiles);
        k_radix_hist<<<n_tiles, SFCI_BLOCK, 0, c->stream>>>(kin, n, dg, hist, n_tiles);
        kt_end(c, t0);
        int t1 = kt_begin(c, K_SCAN, 12.0 * RADIX * n_tiles);
        exclusive_scan_u32(c, hist, (size_t)RADIX * n_tiles, off, nullptr);
        kt_end(c, t1);
        int t2 = kt_begin(c, decltype(dg)::kScatter, 2.0 * (8.0 + sizeof(T)) * n);
        k_radix_scatter<<<n_tiles, SCATTER_THREADS, 0, c->stream>>>(kin, vin, kout, vout, n, dg, off, hist, n_tiles);
        kt_end(c, t2);
        HIPCHK(hipGetLastError());
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    *keys = kin;
    *vals = vin;
}

// Sort n pairs by the low `bits` bits of the key (stable).  (*keys, *vals) hold the
// input; on return they point at the sorted data, which is either the input buffers
// or (ktmp, vtmp).
template <typename T>
void radix_sort_kv(Ctx* c, uint64_t** keys, T** vals, uint64_t* ktmp, T* vtmp, size_t n, int bits) {
    lsd_passes<T>(c, keys, vals, ktmp, vtmp, n, bits, [](int shift) { return DigitPlain{shift}; });
}

void radix_sort_pairs(Ctx* c, uint64_t** keys, double** vals, uint64_t* ktmp, double* vtmp, size_t n, int bits) {
    radix_sort_kv<double>(c, keys, vals, ktmp, vtmp, n, bits);
}

// Group n candidate records by the hb-bit hash of their target (key >> 1): stable LSD
// passes over the hash digits.  Same buffer convention as radix_sort_kv.
template <typename T>
void radix_partition_hash(Ctx* c, uint64_t** keys, T** vals, uint64_t* ktmp, T* vtmp, size_t n, int hb) {
    lsd_passes<T>(c, keys, vals, ktmp, vtmp, n, hb, [hb](int shift) { return DigitHash{shift, hb}; });
}

template void radix_sort_kv<double>(Ctx*, uint64_t**, double**, uint64_t*, double*, size_t, int);
template void radix_sort_kv<float>(Ctx*, uint64_t**, float**, uint64_t*, float*, size_t, int);
template void radix_partition_hash<double>(Ctx*, uint64_t**, double**, uint64_t*, double*, size_t, int);
template void radix_partition_hash<float>(Ctx*, uint64_t**, float**, uint64_t*, float*, size_t, int);

}  // namespace sfci
