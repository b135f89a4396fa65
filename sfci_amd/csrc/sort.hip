// Device-wide exclusive scan and stable LSD radix sort of (u64 key, f64 value) pairs.
//
// Radix sort: 8-bit digits, tiles of 4096 items (256 threads x 16 rounds).  Per pass:
//   hist    - per-tile digit histogram in LDS, written digit-major [256][n_tiles]
//   scan    - exclusive scan of the histogram -> global start of (digit, tile)
//   scatter - stable in-tile ranking: each round of 256 consecutive items is ranked
//             with a 64-lane digit match built from 8 ballots, wave counts are
//             combined in wave order through LDS, so items keep their input order.
// Passes run back to back on the stream with no host synchronization.
#include <vector>

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


// ---- hash partition of candidate records ----------------------------------------------
// Groups records by the hb-bit target hash (hb <= 16) in at most two passes.  Pass 1
// splits the whole list by the top 8 hash bits into 256 regions of fixed capacity
// (expected size + 2% + 64k): each tile ranks its records in LDS, reserves one run per
// digit with an atomic on the region cursor and writes the runs out, so the pass needs
// no histogram.  Pass 2 splits every region by the remaining bits into a dense list
// (histogram, scan, scatter over region-aligned chunks).  Neither pass is stable: records
// of one target are summed exactly, in any order.  A chunk is at most one tile of one
// region; its counter for digit d sits at hbase + d * hstride, so one flat exclusive scan
// gives every output offset of pass 2.
constexpr int PART_TILE = 8192;
constexpr int PART_THREADS = 1024;
constexpr int PART_ITEMS = PART_TILE / PART_THREADS;   // 8
constexpr int HIST_CHUNKS = 16;                        // chunks per histogram workgroup

size_t part_capacity(size_t n) {
    size_t cap1 = ((size_t)((double)n / RADIX * 1.02) + 65536 + 63) & ~(size_t)63;   // chunk starts 64-aligned
    return (size_t)RADIX * cap1;
}

struct Chunks {
    const uint64_t* start;
    const uint32_t* len;
    const uint64_t* hbase;
    const uint32_t* hstride;
    const uint32_t* region;
    size_t n;
};

// chunk table of regions [rstart[r], rstart[r] + rsize[r]); cbase[r] = first chunk of
// region r (n_regions + 1 entries)
__global__ void k_make_chunks(const uint64_t* __restrict__ rstart, const uint64_t* __restrict__ rsize,
                              const uint64_t* __restrict__ cbase, int n_regions, size_t n_chunks, uint32_t tile,
                              uint64_t* __restrict__ start, uint32_t* __restrict__ len, uint64_t* __restrict__ hbase,
                              uint32_t* __restrict__ hstride, uint32_t* __restrict__ region) {
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n_chunks; c += (size_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = n_regions;   // last r with cbase[r] <= c
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (cbase[mid] <= c) lo = mid;
            else hi = mid;
        }
        const uint64_t cl = c - cbase[lo];
        const uint64_t off = cl * tile;
        start[c] = rstart[lo] + off;
        len[c] = (uint32_t)min((uint64_t)tile, rsize[lo] - off);
        hbase[c] = (uint64_t)RADIX * cbase[lo] + cl;
        hstride[c] = (uint32_t)(cbase[lo + 1] - cbase[lo]);
        region[c] = (uint32_t)lo;
    }
}

template <class Digit, typename R>
__global__ __launch_bounds__(SFCI_BLOCK) void k_part_hist(const R* __restrict__ rin, Chunks ch, Digit dg,
                                                          uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[HIST_CHUNKS][RADIX];
    const int tid = threadIdx.x;
    for (int i = tid; i < HIST_CHUNKS * RADIX; i += SFCI_BLOCK) (&h[0][0])[i] = 0;
    __syncthreads();
    const size_t c0 = (size_t)blockIdx.x * HIST_CHUNKS;
    for (int cc = 0; cc < HIST_CHUNKS; ++cc) {
        const size_t c = c0 + cc;
        if (c >= ch.n) break;
        const uint64_t st = ch.start[c];
        const uint32_t len = ch.len[c];
        for (uint32_t i = tid; i < len; i += SFCI_BLOCK) atomicAdd(&h[cc][dg(rec_key(rin[st + i]))], 1u);
    }
    __syncthreads();
    // consecutive threads write consecutive chunks of one digit
    for (int e = tid; e < HIST_CHUNKS * RADIX; e += SFCI_BLOCK) {
        const int d = e / HIST_CHUNKS, cc = e % HIST_CHUNKS;
        const size_t c = c0 + cc;
        if (c < ch.n) hist[ch.hbase[c] + (uint64_t)d * ch.hstride[c]] = h[cc][d];
    }
}

// exclusive scan of one counter per digit held by threads 0..RADIX-1 (all threads call)
__device__ __forceinline__ uint32_t digit_scan(uint32_t h, uint32_t* s_wsum) {
    const int tid = threadIdx.x, lane = tid & (SFCI_WAVE - 1), w = tid / SFCI_WAVE;
    uint32_t x = h;
    if (tid < RADIX) {
        for (int o = 1; o < SFCI_WAVE; o <<= 1) {
            uint32_t y = __shfl_up(x, o, SFCI_WAVE);
            if (lane >= o) x += y;
        }
        if (lane == SFCI_WAVE - 1) s_wsum[w] = x;
    }
    __syncthreads();
    uint32_t pre = 0;
    if (tid < RADIX)
        for (int j = 0; j < w; ++j) pre += s_wsum[j];
    return pre + x - h;
}

// pass 2: chunk c's records go to off[hbase + d * hstride] + (rank among the chunk's digit d)
template <class Digit, typename R>
__global__ __launch_bounds__(PART_THREADS) void k_part_scatter(const R* __restrict__ rin, R* __restrict__ rout,
                                                                Chunks ch, Digit dg, const uint64_t* __restrict__ off,
                                                                const uint32_t* __restrict__ hist) {
    __shared__ R s_rec[PART_TILE];
    __shared__ uint8_t s_dig[PART_TILE];
    __shared__ uint64_t s_gbase[RADIX];
    __shared__ uint32_t s_lstart[RADIX], s_cnt[RADIX];
    __shared__ uint32_t s_wsum[RADIX / SFCI_WAVE];
    const int tid = threadIdx.x;
    const size_t c = blockIdx.x;
    const uint64_t st = ch.start[c];
    const uint32_t len = ch.len[c];
    const uint64_t hb0 = ch.hbase[c];
    const uint32_t hs = ch.hstride[c];
    R x[PART_ITEMS];
#pragma unroll
    for (int r = 0; r < PART_ITEMS; ++r) {
        const uint32_t li = (uint32_t)r * PART_THREADS + tid;
        if (li < len) x[r] = rin[st + li];
    }
    const uint32_t h = tid < RADIX ? hist[hb0 + (uint64_t)tid * hs] : 0u;
    const uint32_t ls = digit_scan(h, s_wsum);
    if (tid < RADIX) {
        s_lstart[tid] = ls;
        s_gbase[tid] = off[hb0 + (uint64_t)tid * hs];
        s_cnt[tid] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PART_ITEMS; ++r) {
        const uint32_t li = (uint32_t)r * PART_THREADS + tid;
        if (li < len) {
            const uint32_t d = dg(rec_key(x[r]));
            const uint32_t p = s_lstart[d] + atomicAdd(&s_cnt[d], 1u);
            s_rec[p] = x[r];
            s_dig[p] = (uint8_t)d;
        }
    }
    __syncthreads();
    for (uint32_t p = tid; p < len; p += PART_THREADS) {
        const uint32_t d = s_dig[p];
        rout[s_gbase[d] + (p - s_lstart[d])] = s_rec[p];
    }
}

// pass 1: tile t = records [t * PART_TILE, ...) of rin; digit d's run is reserved on
// cursor[d] and written only when it ends at or before rend[d]
// dout (optional) receives, at each record's output position, the record's digit under
// dg2: the digit of the next partition pass, so that its histogram reads one byte per record
template <class Digit, typename R>
__global__ __launch_bounds__(PART_THREADS) void k_part_scatter_cur(const R* __restrict__ rin, size_t n,
                                                                    R* __restrict__ rout, Digit dg,
                                                                    unsigned long long* __restrict__ cursor,
                                                                    const uint64_t* __restrict__ rend,
                                                                    uint32_t* __restrict__ flags,
                                                                    uint8_t* __restrict__ dout, Digit dg2) {
    __shared__ R s_rec[PART_TILE];
    __shared__ uint8_t s_dig[PART_TILE];
    __shared__ uint64_t s_gbase[RADIX];
    __shared__ uint32_t s_lstart[RADIX], s_cnt[RADIX], s_ok[RADIX];
    __shared__ uint32_t s_wsum[RADIX / SFCI_WAVE];
    const int tid = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * PART_TILE;
    const uint32_t len = (uint32_t)min((size_t)PART_TILE, n - base);
    R x[PART_ITEMS];
    uint32_t dig[PART_ITEMS], rank[PART_ITEMS];
#pragma unroll
    for (int r = 0; r < PART_ITEMS; ++r) {
        const uint32_t li = (uint32_t)r * PART_THREADS + tid;
        if (li < len) x[r] = rin[base + li];
    }
    if (tid < RADIX) s_cnt[tid] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PART_ITEMS; ++r) {
        const uint32_t li = (uint32_t)r * PART_THREADS + tid;
        dig[r] = 0;
        rank[r] = 0;
        if (li < len) {
            dig[r] = dg(rec_key(x[r]));
            rank[r] = atomicAdd(&s_cnt[dig[r]], 1u);
        }
    }
    __syncthreads();
    const uint32_t h = tid < RADIX ? s_cnt[tid] : 0u;
    const uint32_t ls = digit_scan(h, s_wsum);
    if (tid < RADIX) {
        s_lstart[tid] = ls;
        uint32_t ok = 1;
        if (h) {
            const unsigned long long g = atomicAdd(&cursor[tid], (unsigned long long)h);
            s_gbase[tid] = g;
            if (g + h > rend[tid]) {
                ok = 0;
                atomicOr(flags, FLAG_REGION_FULL);
            }
        }
        s_ok[tid] = ok;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PART_ITEMS; ++r) {
        const uint32_t li = (uint32_t)r * PART_THREADS + tid;
        if (li < len) {
            const uint32_t p = s_lstart[dig[r]] + rank[r];
            s_rec[p] = x[r];
            s_dig[p] = (uint8_t)dig[r];
        }
    }
    __syncthreads();
    for (uint32_t p = tid; p < len; p += PART_THREADS) {
        const uint32_t d = s_dig[p];
        if (s_ok[d]) {
            const uint64_t at = s_gbase[d] + (p - s_lstart[d]);
            const R x = s_rec[p];
            rout[at] = x;
            if (dout) dout[at] = (uint8_t)dg2(rec_key(x));
        }
    }
}

template <typename R, class Digit>
static void part_pass(Ctx* c, const R* rin, R* rout, const Chunks& ch, Digit dg, size_t n) {
    uint32_t* hist = c->scratch_hist.get<uint32_t>((size_t)RADIX * ch.n);
    uint64_t* off = c->scratch_off.get<uint64_t>((size_t)RADIX * ch.n + 1);
    int t0 = kt_begin(c, K_PART_HIST, (double)sizeof(R) * n + 4.0 * RADIX * ch.n);
    k_part_hist<<<(unsigned)((ch.n + HIST_CHUNKS - 1) / HIST_CHUNKS), SFCI_BLOCK, 0, c->stream>>>(rin, ch, dg, hist);
    kt_end(c, t0);
    int t1 = kt_begin(c, K_SCAN, 12.0 * RADIX * ch.n);
    exclusive_scan_u32(c, hist, (size_t)RADIX * ch.n, off, nullptr);
    kt_end(c, t1);
    int t2 = kt_begin(c, K_PART_SCATTER, 2.0 * sizeof(R) * n);
    k_part_scatter<<<(unsigned)ch.n, PART_THREADS, 0, c->stream>>>(rin, rout, ch, dg, off, hist);
    kt_end(c, t2);
    HIPCHK(hipGetLastError());
}

// chunk table of the given regions in c->part_*[slot]
static Chunks make_chunks(Ctx* c, int slot, const std::vector<uint64_t>& rstart, const std::vector<uint64_t>& rsize,
                          uint32_t tile = PART_TILE) {
    const int nr = (int)rstart.size();
    std::vector<uint64_t> meta(3 * nr + 1);
    uint64_t nch = 0;
    for (int r = 0; r < nr; ++r) {
        meta[r] = rstart[r];
        meta[nr + r] = rsize[r];
        meta[2 * nr + r] = nch;
        nch += (rsize[r] + tile - 1) / tile;
    }
    meta[3 * nr] = nch;
    uint64_t* dmeta = c->part_meta[slot].get<uint64_t>(meta.size());
    HIPCHK(hipMemcpyAsync(dmeta, meta.data(), meta.size() * 8, hipMemcpyHostToDevice, c->stream));
    const size_t n_ch = std::max<size_t>(nch, 1);
    uint64_t* st = c->part_start[slot].get<uint64_t>(n_ch);
    uint32_t* ln = c->part_len[slot].get<uint32_t>(n_ch);
    uint64_t* hb = c->part_hbase[slot].get<uint64_t>(n_ch);
    uint32_t* hs = c->part_hstride[slot].get<uint32_t>(n_ch);
    uint32_t* rg = c->part_region[slot].get<uint32_t>(n_ch);
    if (nch) {
        unsigned g = (unsigned)std::min<uint64_t>((nch + SFCI_BLOCK - 1) / SFCI_BLOCK, 4096);
        k_make_chunks<<<g, SFCI_BLOCK, 0, c->stream>>>(dmeta, dmeta + nr, dmeta + 2 * nr, nr, nch, tile, st, ln, hb,
                                                       hs, rg);
    }
    HIPCHK(hipStreamSynchronize(c->stream));   // meta lives on the host stack
    return Chunks{st, ln, hb, hs, rg, (size_t)nch};
}


// ---- second level, chunk-major counters ----------------------------------------------
// cnt[c * RADIX + d]: records of chunk c with digit d.  One workgroup per region sums its
// chunks' counters and sets cursor[r * RADIX + d] to the dense output start of (region r,
// digit d); each chunk then reserves its run of digit d with one atomic on that cursor.
template <class Digit, typename R>
__global__ __launch_bounds__(SFCI_BLOCK) void k_part_hist_cm(const R* __restrict__ rin, Chunks ch, Digit dg,
                                                             uint32_t* __restrict__ cnt) {
    __shared__ uint32_t h[HIST_CHUNKS][RADIX];
    const int tid = threadIdx.x;
    for (int i = tid; i < HIST_CHUNKS * RADIX; i += SFCI_BLOCK) (&h[0][0])[i] = 0;
    __syncthreads();
    const size_t c0 = (size_t)blockIdx.x * HIST_CHUNKS;
    for (int cc = 0; cc < HIST_CHUNKS; ++cc) {
        const size_t c = c0 + cc;
        if (c >= ch.n) break;
        const uint64_t st = ch.start[c];
        const uint32_t len = ch.len[c];
        for (uint32_t i = tid; i < len; i += SFCI_BLOCK) atomicAdd(&h[cc][dg(rec_key(rin[st + i]))], 1u);
    }
    __syncthreads();
    for (int e = tid; e < HIST_CHUNKS * RADIX; e += SFCI_BLOCK) {
        const size_t c = c0 + e / RADIX;
        if (c < ch.n) cnt[c * RADIX + (e % RADIX)] = (&h[0][0])[e];
    }
}

// the same counters from the digit bytes written by the previous pass (chunk starts are
// 64-aligned: four digits per 32-bit load)
__global__ __launch_bounds__(SFCI_BLOCK) void k_part_hist_bytes(const uint8_t* __restrict__ din, Chunks ch,
                                                                uint32_t* __restrict__ cnt) {
    __shared__ uint32_t h[HIST_CHUNKS][RADIX];
    const int tid = threadIdx.x;
    for (int i = tid; i < HIST_CHUNKS * RADIX; i += SFCI_BLOCK) (&h[0][0])[i] = 0;
    __syncthreads();
    const size_t c0 = (size_t)blockIdx.x * HIST_CHUNKS;
    for (int cc = 0; cc < HIST_CHUNKS; ++cc) {
        const size_t c = c0 + cc;
        if (c >= ch.n) break;
        const uint64_t st = ch.start[c];
        const uint32_t len = ch.len[c];
        const uint32_t* w = (const uint32_t*)(din + st);
        // HIST_UNROLL words per thread loaded before any is counted
        constexpr int HIST_UNROLL = 8;
        const uint32_t nw = (len + 3) / 4;
        for (uint32_t i0 = 0; i0 < nw; i0 += HIST_UNROLL * SFCI_BLOCK) {
            uint32_t x[HIST_UNROLL];
#pragma unroll
            for (int u = 0; u < HIST_UNROLL; ++u) {
                const uint32_t i = i0 + (uint32_t)u * SFCI_BLOCK + tid;
                x[u] = i < nw ? w[i] : 0u;
            }
#pragma unroll
            for (int u = 0; u < HIST_UNROLL; ++u) {
                const uint32_t i = i0 + (uint32_t)u * SFCI_BLOCK + tid;
                if (i < nw) {
                    const uint32_t m = min(4u, len - 4 * i);
                    for (uint32_t e = 0; e < m; ++e) atomicAdd(&h[cc][(x[u] >> (8 * e)) & 255u], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < HIST_CHUNKS * RADIX; e += SFCI_BLOCK) {
        const size_t c = c0 + e / RADIX;
        if (c < ch.n) cnt[c * RADIX + (e % RADIX)] = (&h[0][0])[e];
    }
}

constexpr int COLSCAN_UNROLL = 8;

__global__ __launch_bounds__(RADIX) void k_region_colscan(const uint32_t* __restrict__ cnt,
                                                          const uint64_t* __restrict__ cbase,
                                                          const uint64_t* __restrict__ dense_start,
                                                          unsigned long long* __restrict__ cursor) {
    __shared__ uint32_t s_wsum[RADIX / SFCI_WAVE];
    const int r = blockIdx.x, d = threadIdx.x;
    const uint64_t c0 = cbase[r], c1 = cbase[r + 1];
    uint32_t run = 0;
    uint64_t cc = c0;
    for (; cc + COLSCAN_UNROLL <= c1; cc += COLSCAN_UNROLL) {
        uint32_t v[COLSCAN_UNROLL];
#pragma unroll
        for (int u = 0; u < COLSCAN_UNROLL; ++u) v[u] = cnt[(cc + u) * RADIX + d];
#pragma unroll
        for (int u = 0; u < COLSCAN_UNROLL; ++u) run += v[u];
    }
    for (; cc < c1; ++cc) run += cnt[cc * RADIX + d];
    const uint32_t pre = digit_scan(run, s_wsum);
    cursor[(uint64_t)r * RADIX + d] = dense_start[r] + pre;
}

// a chunk reserves one run per digit on its region's (digit) cursor
template <class Digit, typename R>
__global__ __launch_bounds__(PART_THREADS) void k_part_scatter_cm(const R* __restrict__ rin, R* __restrict__ rout,
                                                                   Chunks ch, Digit dg, const uint32_t* __restrict__ cnt,
                                                                   unsigned long long* __restrict__ cursor) {
    __shared__ R s_rec[PART_TILE];
    __shared__ uint8_t s_dig[PART_TILE];
    __shared__ uint64_t s_gbase[RADIX];
    __shared__ uint32_t s_lstart[RADIX], s_cnt[RADIX];
    __shared__ uint32_t s_wsum[RADIX / SFCI_WAVE];
    const int tid = threadIdx.x;
    const size_t c = blockIdx.x;
    const uint64_t st = ch.start[c];
    const uint32_t len = ch.len[c];
    const uint32_t rg = ch.region[c];
    R x[PART_ITEMS];
#pragma unroll
    for (int r = 0; r < PART_ITEMS; ++r) {
        const uint32_t li = (uint32_t)r * PART_THREADS + tid;
        if (li < len) x[r] = rin[st + li];
    }
    const uint32_t h = tid < RADIX ? cnt[c * RADIX + tid] : 0u;
    const uint32_t ls = digit_scan(h, s_wsum);
    if (tid < RADIX) {
        s_lstart[tid] = ls;
        s_gbase[tid] = h ? atomicAdd(&cursor[(uint64_t)rg * RADIX + tid], (unsigned long long)h) : 0ull;
        s_cnt[tid] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PART_ITEMS; ++r) {
        const uint32_t li = (uint32_t)r * PART_THREADS + tid;
        if (li < len) {
            const uint32_t d = dg(rec_key(x[r]));
            const uint32_t p = s_lstart[d] + atomicAdd(&s_cnt[d], 1u);
            s_rec[p] = x[r];
            s_dig[p] = (uint8_t)d;
        }
    }
    __syncthreads();
    for (uint32_t p = tid; p < len; p += PART_THREADS) {
        const uint32_t d = s_dig[p];
        rout[s_gbase[d] + (p - s_lstart[d])] = s_rec[p];
    }
}

// regions given by (rstart, rsize) with dense output starts dense_start: rin -> rout
template <typename R, class Digit>
static void part_pass_regions(Ctx* c, const R* rin, R* rout, const std::vector<uint64_t>& rstart,
                              const std::vector<uint64_t>& rsize, Digit dg, size_t n, const uint8_t* din = nullptr) {
    const int nr = (int)rstart.size();
    Chunks ch = make_chunks(c, 0, rstart, rsize);
    // meta of slot 0 holds cbase at [2 nr, 3 nr]; dense starts go to slot 1
    const uint64_t* cbase = (const uint64_t*)c->part_meta[0].p + 2 * nr;
    std::vector<uint64_t> ds(nr);
    uint64_t acc = 0;
    for (int r = 0; r < nr; ++r) {
        ds[r] = acc;
        acc += rsize[r];
    }
    uint64_t* dds = c->part_meta[1].get<uint64_t>(nr);
    HIPCHK(hipMemcpyAsync(dds, ds.data(), nr * 8, hipMemcpyHostToDevice, c->stream));
    uint32_t* cnt = c->scratch_hist.get<uint32_t>((size_t)RADIX * std::max<size_t>(ch.n, 1));
    unsigned long long* cursor = (unsigned long long*)c->part_base.get<uint64_t>((size_t)RADIX * nr);
    int t0 = kt_begin(c, K_PART_HIST, (din ? 1.0 : (double)sizeof(R)) * n + 4.0 * RADIX * ch.n);
    if (din)
        k_part_hist_bytes<<<(unsigned)((ch.n + HIST_CHUNKS - 1) / HIST_CHUNKS), SFCI_BLOCK, 0, c->stream>>>(din, ch,
                                                                                                           cnt);
    else
        k_part_hist_cm<<<(unsigned)((ch.n + HIST_CHUNKS - 1) / HIST_CHUNKS), SFCI_BLOCK, 0, c->stream>>>(rin, ch, dg,
                                                                                                        cnt);
    kt_end(c, t0);
    int t1 = kt_begin(c, K_SCAN, 4.0 * RADIX * ch.n);
    k_region_colscan<<<nr, RADIX, 0, c->stream>>>(cnt, cbase, dds, cursor);
    kt_end(c, t1);
    int t2 = kt_begin(c, K_PART_SCATTER, 2.0 * sizeof(R) * n);
    k_part_scatter_cm<<<(unsigned)ch.n, PART_THREADS, 0, c->stream>>>(rin, rout, ch, dg, cnt, cursor);
    kt_end(c, t2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));   // ds lives on the host stack
}

template <typename R>
R* hash_partition(Ctx* c, R* A, R* B, size_t n, int hb) {
    if (hb <= 0 || hb > 16) throw HipError("hash_partition: 1..16 hash bits", -1);
    if (hb <= 8) {
        Chunks ch = make_chunks(c, 0, {0}, {(uint64_t)n});
        part_pass<R>(c, A, B, ch, DigitHash{0, hb}, n);
        return B;
    }
    // pass 1: top 8 hash bits, fixed-capacity regions of B
    const size_t cap1 = part_capacity(n) / RADIX;
    std::vector<uint64_t> cur(2 * RADIX);
    for (int d = 0; d < RADIX; ++d) {
        cur[d] = (uint64_t)d * cap1;
        cur[RADIX + d] = (uint64_t)(d + 1) * cap1;
    }
    uint64_t* dcur = c->part_meta[1].get<uint64_t>(2 * RADIX);
    HIPCHK(hipMemcpyAsync(dcur, cur.data(), 2 * RADIX * 8, hipMemcpyHostToDevice, c->stream));
    const size_t n_tiles = (n + PART_TILE - 1) / PART_TILE;
    uint8_t* dig = c->part_dig.get<uint8_t>(part_capacity(n));
    int t1 = kt_begin(c, K_PART_SCATTER1, (2.0 * sizeof(R) + 1.0) * n);
    k_part_scatter_cur<<<(unsigned)n_tiles, PART_THREADS, 0, c->stream>>>(
        A, n, B, DigitHash{hb - 8, hb}, (unsigned long long*)dcur, dcur + RADIX, c->d_flags, dig, DigitHash{0, hb});
    kt_end(c, t1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(cur.data(), dcur, RADIX * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    uint32_t fl = flags_read_clear(c);
    if (fl & ~(uint32_t)FLAG_REGION_FULL) atomic_flags_restore(c, fl & ~(uint32_t)FLAG_REGION_FULL);
    if (fl & FLAG_REGION_FULL) return nullptr;
    // pass 2: each region by the remaining hash bits (the pass-1 bits above them are
    // constant inside a region), into a dense list in A
    std::vector<uint64_t> rstart(RADIX), rsize(RADIX);
    for (int d = 0; d < RADIX; ++d) {
        rstart[d] = (uint64_t)d * cap1;
        rsize[d] = cur[d] - rstart[d];
    }
    part_pass_regions<R>(c, B, A, rstart, rsize, DigitHash{0, hb}, n, dig);
    return A;
}

// ---- hash partition of compact records ----------------------------------------------
// The same two passes for compact records, whose hash bucket (hb bits, 8 < hb <= 16) is
// not in the key: pass 1 takes its digit (the bucket's top 8 bits) from the 16-bit bucket
// array written by the generation and writes the bucket's low hb - 8 bits as the digit byte
// of pass 2, which reads both its counts and its scatter digit from those bytes.  The
// bucket bounds come from pass 2's per-(region, digit) output starts.
// records per tile of these passes: 8192 (4096-record tiles, two workgroups per CU, measured
// slower), 4096 for records above 16 bytes so that a tile fits in LDS
template <typename R>
constexpr int ptile() {
    return sizeof(R) > 16 ? PART_TILE / 2 : PART_TILE;
}
template <typename R>
constexpr int pthreads() {
    return ptile<R>() / PART_ITEMS;
}

// A tile of records in registers.
template <typename R>
struct TileRegs {
    static constexpr int NW = sizeof(R) / 4;   // records held as 32-bit words
    uint32_t x[PART_ITEMS][NW];
    uint16_t d[PART_ITEMS];                    // hash bucket
};

template <typename R>
__device__ __forceinline__ void tile_load(const R* __restrict__ rin, const uint16_t* __restrict__ bkt, size_t base,
                                          uint32_t len, TileRegs<R>& t) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < PART_ITEMS; ++r) {
        const uint32_t li = (uint32_t)r * pthreads<R>() + tid;
        if (li < len) {
            const uint32_t* src = (const uint32_t*)(rin + base + li);
#pragma unroll
            for (int q = 0; q < TileRegs<R>::NW; ++q) t.x[r][q] = src[q];
            t.d[r] = bkt[base + li];
        }
    }
}

// First level: one persistent workgroup per CU (a tile's LDS staging takes most of the CU's
// LDS, so one workgroup runs per CU): while a tile is ranked and written out, the next
// tile's records are already loaded into registers, so the reads overlap the writes.
constexpr size_t PART_PERSIST = 256;
template <typename R>
__global__ __launch_bounds__(pthreads<R>()) void k_part_scatter_cur_c(const R* __restrict__ rin,
                                                                      const uint16_t* __restrict__ bkt, size_t n,
                                                                      R* __restrict__ rout, int sh1,
                                                                      unsigned long long* __restrict__ cursor,
                                                                      const uint64_t* __restrict__ rend,
                                                                      uint32_t* __restrict__ flags,
                                                                      uint8_t* __restrict__ dout,
                                                                      unsigned long long* __restrict__ tile_next) {
    __shared__ R s_rec[ptile<R>()];
    __shared__ uint16_t s_bk[ptile<R>()];
    __shared__ uint64_t s_gbase[RADIX];
    __shared__ uint32_t s_lstart[RADIX], s_cnt[RADIX], s_ok[RADIX];
    __shared__ uint32_t s_wsum[RADIX / SFCI_WAVE];
    const int tid = threadIdx.x;
    const uint32_t m2 = (1u << sh1) - 1u;
    const size_t n_tiles = (n + ptile<R>() - 1) / ptile<R>();
    constexpr int NW = TileRegs<R>::NW;
    TileRegs<R> cur, nxt;
    uint32_t rank[PART_ITEMS];
    // tiles are taken from a global counter (a workgroup's first tile is its index), so a
    // workgroup that runs slower takes fewer tiles; the tile after next is claimed while this
    // one is ranked, and published in s_tn at the barrier before the write-out
    __shared__ unsigned long long s_tn;
    size_t t = blockIdx.x;
    if (tid == 0) s_tn = gridDim.x + atomicAdd(tile_next, 1ull);
    if (t < n_tiles) tile_load(rin, bkt, t * ptile<R>(), (uint32_t)min((size_t)ptile<R>(), n - t * ptile<R>()), cur);
    while (t < n_tiles) {
        const size_t base = t * ptile<R>();
        const uint32_t len = (uint32_t)min((size_t)ptile<R>(), n - base);
        if (tid < RADIX) s_cnt[tid] = 0;
        __syncthreads();
        const size_t tn = (size_t)s_tn;   // this workgroup's next tile
        unsigned long long tnn = 0;
        if (tid == 0) tnn = gridDim.x + atomicAdd(tile_next, 1ull);   // and the one after
#pragma unroll
        for (int r = 0; r < PART_ITEMS; ++r) {
            const uint32_t li = (uint32_t)r * pthreads<R>() + tid;
            rank[r] = 0;
            if (li < len) rank[r] = atomicAdd(&s_cnt[(uint32_t)cur.d[r] >> sh1], 1u);
        }
        __syncthreads();
        const uint32_t h = tid < RADIX ? s_cnt[tid] : 0u;
        const uint32_t ls = digit_scan(h, s_wsum);
        if (tid < RADIX) {
            s_lstart[tid] = ls;
            uint32_t ok = 1;
            if (h) {
                const unsigned long long g = atomicAdd(&cursor[tid], (unsigned long long)h);
                s_gbase[tid] = g;
                if (g + h > rend[tid]) {
                    ok = 0;
                    atomicOr(flags, FLAG_REGION_FULL);
                }
            }
            s_ok[tid] = ok;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < PART_ITEMS; ++r) {
            const uint32_t li = (uint32_t)r * pthreads<R>() + tid;
            if (li < len) {
                const uint32_t p = s_lstart[(uint32_t)cur.d[r] >> sh1] + rank[r];
                uint32_t* dst = (uint32_t*)(s_rec + p);
#pragma unroll
                for (int q = 0; q < NW; ++q) dst[q] = cur.x[r][q];
                s_bk[p] = cur.d[r];
            }
        }
        if (tid == 0) s_tn = tnn;   // read by every thread after the next iteration's first barrier
        __syncthreads();
        // the next tile's loads go out before this tile's stores
        if (tn < n_tiles)
            tile_load(rin, bkt, tn * ptile<R>(), (uint32_t)min((size_t)ptile<R>(), n - tn * ptile<R>()), nxt);
        for (uint32_t p = tid; p < len; p += pthreads<R>()) {
            const uint32_t b = s_bk[p], d = b >> sh1;
            if (s_ok[d]) {
                const uint64_t at = s_gbase[d] + (p - s_lstart[d]);
                rout[at] = s_rec[p];
                dout[at] = (uint8_t)(b & m2);
            }
        }
        cur = nxt;
        t = tn;
        __syncthreads();   // LDS reuse by the next tile
    }
}

// second level: a chunk reserves one run per digit (from the digit bytes) on its region's
// (digit) cursor
template <typename R>
__global__ __launch_bounds__(pthreads<R>()) void k_part_scatter_cm_c(const R* __restrict__ rin,
                                                                     const uint8_t* __restrict__ din,
                                                                     R* __restrict__ rout, Chunks ch,
                                                                     const uint32_t* __restrict__ cnt,
                                                                     unsigned long long* __restrict__ cursor) {
    __shared__ R s_rec[ptile<R>()];
    __shared__ uint8_t s_dig[ptile<R>()];
    __shared__ uint64_t s_gbase[RADIX];
    __shared__ uint32_t s_lstart[RADIX], s_cnt[RADIX];
    __shared__ uint32_t s_wsum[RADIX / SFCI_WAVE];
    const int tid = threadIdx.x;
    const size_t c = blockIdx.x;
    const uint64_t st = ch.start[c];
    const uint32_t len = ch.len[c];
    const uint32_t rg = ch.region[c];
    constexpr int NW = sizeof(R) / 4;   // records held as 32-bit words (kept in registers)
    uint32_t x[PART_ITEMS][NW];
    uint32_t dg[PART_ITEMS];
#pragma unroll
    for (int r = 0; r < PART_ITEMS; ++r) {
        const uint32_t li = (uint32_t)r * pthreads<R>() + tid;
        if (li < len) {
            const uint32_t* src = (const uint32_t*)(rin + st + li);
#pragma unroll
            for (int q = 0; q < NW; ++q) x[r][q] = src[q];
            dg[r] = din[st + li];
        }
    }
    const uint32_t h = tid < RADIX ? cnt[c * RADIX + tid] : 0u;
    const uint32_t ls = digit_scan(h, s_wsum);
    if (tid < RADIX) {
        s_lstart[tid] = ls;
        s_gbase[tid] = h ? atomicAdd(&cursor[(uint64_t)rg * RADIX + tid], (unsigned long long)h) : 0ull;
        s_cnt[tid] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PART_ITEMS; ++r) {
        const uint32_t li = (uint32_t)r * pthreads<R>() + tid;
        if (li < len) {
            const uint32_t p = s_lstart[dg[r]] + atomicAdd(&s_cnt[dg[r]], 1u);
            uint32_t* dst = (uint32_t*)(s_rec + p);
#pragma unroll
            for (int q = 0; q < NW; ++q) dst[q] = x[r][q];
            s_dig[p] = (uint8_t)dg[r];
        }
    }
    __syncthreads();
    for (uint32_t p = tid; p < len; p += pthreads<R>()) {
        const uint32_t d = s_dig[p];
        rout[s_gbase[d] + (p - s_lstart[d])] = s_rec[p];
    }
}

// bucket b = (region r, digit d) with r = b >> sh1, d = b & (2^sh1 - 1): its records start
// at the saved pass-2 output start of (r, d) and end where bucket b + 1 starts
__global__ void k_bucket_bounds_c(const unsigned long long* __restrict__ start_rd, int sh1, uint64_t n_buckets,
                                  size_t n, uint64_t* __restrict__ bstart, uint64_t* __restrict__ bend) {
    const uint64_t m2 = ((uint64_t)1 << sh1) - 1;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < n_buckets;
         b += (uint64_t)gridDim.x * blockDim.x) {
        bstart[b] = start_rd[(b >> sh1) * RADIX + (b & m2)];
        const uint64_t b1 = b + 1;
        bend[b] = b1 < n_buckets ? start_rd[(b1 >> sh1) * RADIX + (b1 & m2)] : n;
    }
}

template <typename R>
bool hash_partition_c(Ctx* c, R* A, const uint16_t* bkt, R* B, size_t n, int hb, uint64_t* bstart,
                      uint64_t* bend) {
    if (hb <= 8 || hb > 16) throw HipError("hash_partition_c: 9..16 hash bits", -1);
    const int sh1 = hb - 8;
    // pass 1: the bucket's top 8 bits, fixed-capacity regions of B
    const size_t cap1 = part_capacity(n) / RADIX;
    // region cursors, region ends, and the tile counter
    std::vector<uint64_t> cur(2 * RADIX + 1, 0);
    for (int d = 0; d < RADIX; ++d) {
        cur[d] = (uint64_t)d * cap1;
        cur[RADIX + d] = (uint64_t)(d + 1) * cap1;
    }
    uint64_t* dcur = c->part_meta[1].get<uint64_t>(2 * RADIX + 1);
    HIPCHK(hipMemcpyAsync(dcur, cur.data(), (2 * RADIX + 1) * 8, hipMemcpyHostToDevice, c->stream));
    const size_t n_tiles = (n + ptile<R>() - 1) / ptile<R>();
    uint8_t* dig = c->part_dig.get<uint8_t>(part_capacity(n));
    int t1 = kt_begin(c, K_PART_SCATTER1, (2.0 * sizeof(R) + 3.0) * n);
    k_part_scatter_cur_c<R><<<(unsigned)std::min<size_t>(n_tiles, PART_PERSIST), pthreads<R>(), 0, c->stream>>>(
        A, bkt, n, B, sh1, (unsigned long long*)dcur, dcur + RADIX, c->d_flags, dig,
        (unsigned long long*)dcur + 2 * RADIX);
    kt_end(c, t1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(cur.data(), dcur, RADIX * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    uint32_t fl = flags_read_clear(c);
    if (fl & ~(uint32_t)FLAG_REGION_FULL) atomic_flags_restore(c, fl & ~(uint32_t)FLAG_REGION_FULL);
    if (fl & FLAG_REGION_FULL) return false;
    // pass 2: each region by the bucket's low bits (from the digit bytes), dense into A
    std::vector<uint64_t> rstart(RADIX), rsize(RADIX);
    for (int d = 0; d < RADIX; ++d) {
        rstart[d] = (uint64_t)d * cap1;
        rsize[d] = cur[d] - rstart[d];
    }
    const int nr = RADIX;
    Chunks ch = make_chunks(c, 0, rstart, rsize, ptile<R>());
    const uint64_t* cbase = (const uint64_t*)c->part_meta[0].p + 2 * nr;
    std::vector<uint64_t> ds(nr);
    uint64_t acc = 0;
    for (int r = 0; r < nr; ++r) {
        ds[r] = acc;
        acc += rsize[r];
    }
    uint64_t* dds = c->part_meta[1].get<uint64_t>(nr);
    HIPCHK(hipMemcpyAsync(dds, ds.data(), nr * 8, hipMemcpyHostToDevice, c->stream));
    uint32_t* cnt = c->scratch_hist.get<uint32_t>((size_t)RADIX * std::max<size_t>(ch.n, 1));
    unsigned long long* cursor = (unsigned long long*)c->part_base.get<uint64_t>((size_t)2 * RADIX * nr);
    unsigned long long* start_rd = cursor + (size_t)RADIX * nr;
    int t0 = kt_begin(c, K_PART_HIST, 1.0 * n + 4.0 * RADIX * ch.n);
    k_part_hist_bytes<<<(unsigned)((ch.n + HIST_CHUNKS - 1) / HIST_CHUNKS), SFCI_BLOCK, 0, c->stream>>>(dig, ch, cnt);
    kt_end(c, t0);
    int t2 = kt_begin(c, K_SCAN, 4.0 * RADIX * ch.n);
    k_region_colscan<<<nr, RADIX, 0, c->stream>>>(cnt, cbase, dds, cursor);
    kt_end(c, t2);
    HIPCHK(hipMemcpyAsync(start_rd, cursor, (size_t)RADIX * nr * 8, hipMemcpyDeviceToDevice, c->stream));
    int t3 = kt_begin(c, K_PART_SCATTER, (2.0 * sizeof(R) + 1.0) * n);
    k_part_scatter_cm_c<R><<<(unsigned)ch.n, pthreads<R>(), 0, c->stream>>>(B, dig, A, ch, cnt, cursor);
    kt_end(c, t3);
    const uint64_t nb = (uint64_t)1 << hb;
    int tb = kt_begin(c, K_BUCKET_BOUNDS, 16.0 * nb);
    k_bucket_bounds_c<<<(unsigned)((nb + SFCI_BLOCK - 1) / SFCI_BLOCK), SFCI_BLOCK, 0, c->stream>>>(
        start_rd, sh1, nb, n, bstart, bend);
    kt_end(c, tb);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));   // ds lives on the host stack
    return true;
}

template bool hash_partition_c<RecC64>(Ctx*, RecC64*, const uint16_t*, RecC64*, size_t, int, uint64_t*, uint64_t*);
template bool hash_partition_c<RecC32>(Ctx*, RecC32*, const uint16_t*, RecC32*, size_t, int, uint64_t*, uint64_t*);
template bool hash_partition_c<RecW>(Ctx*, RecW*, const uint16_t*, RecW*, size_t, int, uint64_t*, uint64_t*);
template Rec64* hash_partition<Rec64>(Ctx*, Rec64*, Rec64*, size_t, int);
template Rec32* hash_partition<Rec32>(Ctx*, Rec32*, Rec32*, size_t, int);

template void radix_sort_kv<double>(Ctx*, uint64_t**, double**, uint64_t*, double*, size_t, int);
template void radix_sort_kv<float>(Ctx*, uint64_t**, float**, uint64_t*, float*, size_t, int);
template void radix_sort_kv<uint64_t>(Ctx*, uint64_t**, uint64_t**, uint64_t*, uint64_t*, size_t, int);
template void radix_partition_hash<double>(Ctx*, uint64_t**, double**, uint64_t*, double*, size_t, int);
template void radix_partition_hash<float>(Ctx*, uint64_t**, float**, uint64_t*, float*, size_t, int);

}  // namespace sfci
