// Screening network on the GPU: a 66 -> 256 -> 256 -> 1 MLP with SiLU activations that
// predicts log|c| of a candidate determinant from its features (occupation +-1 of 64 spin
// orbitals, excitation rank of the connection, plog2|H_ij c_j|: exponent + mantissa
// fraction of the FP32 value, exact to compute anywhere).  No reference counterpart:
// the reference keeps every candidate (src/functions.cpp:623-639).
//
// Arithmetic: weights and activations are BF16 (round to nearest even), products are
// accumulated in FP32 by the gfx950 matrix cores (v_mfma_f32_32x32x16_bf16), biases and
// the master copy of every parameter are FP32.
//   z1 = X W1^T + b1,  h1 = bf16(silu(z1)),  z2 = h1 W2^T + b2,  h2 = bf16(silu(z2)),
//   y  = h2 . w3 + b3  (w3 rounded to BF16, the dot in FP32)
// Training minimises the mean squared error with Adam; the backward pass rounds dz2 and
// dz1 to BF16 (the operands of its matrix products), everything else stays FP32.
//
// Matrix-core fragments (one wavefront, 32 x 32 x 16): lane l = 32 g + i holds row i of
// the A tile and row i of the B^T tile, each with the 8 consecutive k values 8g..8g+7;
// the accumulator's element r of lane l is C[8 (r / 4) + 4 g + r % 4][i].  A and B use
// the same k assignment, so the product does not depend on how the hardware pairs them.
#include <cmath>

#include "../../include/sfci_gpu.h"
#include "internal.h"

using namespace sfci;

namespace sfci {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t bf16_t;

constexpr int IN = MLP_IN, INP = MLP_INP, HID = MLP_HID;
constexpr int XT_ROWS = 96;          // feature rows of the transposed batch (80 padded to 3 x 32)
constexpr int FW_TM = 64;            // rows per tile of the fused forward
constexpr int FW_MT = FW_TM / 32;     // 32-row MFMA tiles per tile
constexpr int FW_THREADS = 512;       // eight wavefronts, two per SIMD
constexpr int FW_FPT = FW_THREADS / FW_TM;   // threads per row building features
constexpr int FW_FCOLS = 80 / FW_FPT;        // feature columns per thread
constexpr int FW_XS = 88;            // LDS row stride (BF16) of the feature tile
constexpr int FW_HS = 264;           // LDS row stride (BF16) of the hidden tile

__host__ __device__ __forceinline__ bf16_t f2bf(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((u >> 16) | 0x40u);   // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(bf16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
// BF16 rounding (nearest even) in one instruction on gfx950 (v_cvt_pk_bf16_f32); the same
// result as f2bf for every finite value
__device__ __forceinline__ bf16_t f2bf_hw(float f) {
    const __bf16 h = (__bf16)f;
    bf16_t u;
    __builtin_memcpy(&u, &h, 2);
    return u;
}
// sigmoid with the hardware exp2 and reciprocal (about 1 ulp; the activations are then
// rounded to BF16, 2^-9 relative)
__device__ __forceinline__ float sigm(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ bf16x8 ld8(const bf16_t* p) { return *(const bf16x8*)p; }
__device__ __forceinline__ f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Device view of the parameters.
struct MlpDev {
    const bf16_t* w1t;   // [HID][INP], columns >= IN zero
    const float* b1;     // [HID]
    const bf16_t* w2t;   // [HID][HID]: row n = hidden-2 unit, column j = hidden-1 unit
    const float* b2;     // [HID]
    const float* w3r;    // [HID] BF16-rounded output weights, as FP32
    const float* b3;     // [1]
    float x_shift, x_scale;   // feature 65 = (plog2|v| - x_shift) * x_scale
    // candidate scores (records in): out_mul * output + mag_mul * plog2|v|; 1 and 0 give the
    // network's output itself
    float out_mul, mag_mul;
};

// Optional activations saved by the training forward (row-major [n][HID]).
struct FwdSave {
    float* z1;
    bf16_t* h1;
    float* z2;
    bf16_t* h2;
};

// Feature `col` of a candidate: +-1 occupation for spin orbitals < n_so (0 above), the
// excitation rank, the scaled plog2|value|, zero padding up to INP.
__device__ __forceinline__ float feature(int col, uint64_t det, int n_so, uint32_t cls, float lv, const MlpDev& P) {
    if (col < 64) return col < n_so ? (((det >> col) & 1ull) ? 1.0f : -1.0f) : 0.0f;
    if (col == 64) return (float)cls;
    if (col == 65) return (lv - P.x_shift) * P.x_scale;
    return 0.0f;
}

enum { IN_X = 0, IN_REC = 1, IN_SAMPLE = 2 };

// Fused three-layer forward, persistent over tiles of FW_TM rows, one workgroup of eight
// wavefronts (two per SIMD, so that one wavefront's matrix-core work overlaps another's
// epilogue).  Wavefront w owns hidden columns [32 w, 32 w + 32) of both layers and keeps
// their weight fragments (B operands) in registers for the whole launch, so the weights are
// read from memory once per workgroup.  Per tile: features -> LDS (built from inputs loaded
// during the previous tile); layer 1 (A from LDS) -> bias, SiLU, BF16 -> LDS; layer 2 (A from
// LDS) -> bias, SiLU, BF16 -> LDS; the output dot over each row is summed by four threads
// over 64 columns each and the four partials are added in a fixed order.  Every output
// element is the same chain of MFMA k-steps whatever the tile shape.
// Inputs: IN_X, BF16 features [n][INP]; IN_REC, screened records; IN_SAMPLE, samples.
template <int MODE>
__global__ __launch_bounds__(FW_THREADS) void k_mlp_fwd(MlpDev P, int n_so, const void* __restrict__ in, size_t n,
                                                        float* __restrict__ y, FwdSave sv) {
    __shared__ __attribute__((aligned(16))) bf16_t sx[FW_TM * FW_XS];
    __shared__ __attribute__((aligned(16))) bf16_t sh[FW_TM * FW_HS];
    __shared__ float sw3[HID];
    __shared__ float ypart[4][FW_TM];
    static_assert(FW_FCOLS * FW_FPT == 80 && FW_FCOLS % 2 == 0, "feature columns split evenly in pairs");
    static_assert(4 * FW_TM <= FW_THREADS, "the output dot uses four threads per row");
    static_assert(FW_THREADS / 64 * 32 == HID, "each wavefront owns 32 hidden columns");
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, li = lane & 31, g = lane >> 5;
    if (tid < HID) sw3[tid] = P.w3r[tid];
    bf16x8 w1f[INP / 16], w2f[HID / 16];
    const int col = w * 32 + li;
    for (int ks = 0; ks < INP / 16; ++ks) w1f[ks] = ld8(P.w1t + (size_t)col * INP + ks * 16 + g * 8);
    for (int ks = 0; ks < HID / 16; ++ks) w2f[ks] = ld8(P.w2t + (size_t)col * HID + ks * 16 + g * 8);
    const float b1c = P.b1[col], b2c = P.b2[col];
    const float b3 = P.b3[0];
    const size_t n_tiles = (n + FW_TM - 1) / FW_TM;
    // this thread's feature inputs: row fr of the tile, columns [FW_FCOLS part, + FW_FCOLS)
    const int fr = tid / FW_FPT, part = tid % FW_FPT;
    uint64_t p_det = 0;
    uint32_t p_cls = 0;
    float p_lv = 0.0f;
    uint32_t p_x[FW_FCOLS / 2] = {};   // IN_X: BF16 values, two per word
    auto fetch = [&](size_t tile) {
        const size_t gr = tile * FW_TM + fr;
        if (tile >= n_tiles || gr >= n) return;
        if constexpr (MODE == IN_X) {
            const uint32_t* src = (const uint32_t*)((const bf16_t*)in + gr * INP + part * FW_FCOLS);
            for (int c2 = 0; c2 < FW_FCOLS / 2; ++c2) p_x[c2] = src[c2];
        } else if constexpr (MODE == IN_REC) {
            const RecS rr = ((const RecS*)in)[gr];
            p_det = rr.det;
            p_cls = rr.cls;
            p_lv = plog2(rr.val);
        } else {
            const sfci_screen_sample_t s = ((const sfci_screen_sample_t*)in)[gr];
            p_det = s.det;
            p_cls = s.cls;
            p_lv = s.logv;
        }
    };
    fetch(blockIdx.x);
    for (size_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const size_t row0 = tile * FW_TM;
        // features of this tile from the inputs fetched earlier; then the next tile's inputs
        {
            bf16_t* dst = sx + fr * FW_XS + part * FW_FCOLS;
            if (row0 + fr >= n) {
                for (int c2 = 0; c2 < FW_FCOLS; ++c2) dst[c2] = 0;
            } else if constexpr (MODE == IN_X) {
                uint32_t* d32 = (uint32_t*)dst;
                for (int c2 = 0; c2 < FW_FCOLS / 2; ++c2) d32[c2] = p_x[c2];
            } else {
                for (int c2 = 0; c2 < FW_FCOLS; ++c2)
                    dst[c2] = f2bf_hw(feature(part * FW_FCOLS + c2, p_det, n_so, p_cls, p_lv, P));
            }
        }
        fetch(tile + gridDim.x);
        __syncthreads();
        // layer 1
        f32x16 acc[FW_MT];
        for (int a = 0; a < FW_MT; ++a)
            for (int r = 0; r < 16; ++r) acc[a][r] = 0.0f;
#pragma unroll
        for (int ks = 0; ks < INP / 16; ++ks) {
            bf16x8 av[FW_MT];
#pragma unroll
            for (int mt = 0; mt < FW_MT; ++mt) av[mt] = ld8(sx + (mt * 32 + li) * FW_XS + ks * 16 + g * 8);
#pragma unroll
            for (int mt = 0; mt < FW_MT; ++mt) acc[mt] = mfma32(av[mt], w1f[ks], acc[mt]);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int mt = 0; mt < FW_MT; ++mt) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = mt * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
                const float z = acc[mt][r] + b1c;
                const bf16_t h = f2bf_hw(silu(z));
                sh[row * FW_HS + col] = h;
                if (sv.z1 && row0 + row < n) {
                    sv.z1[(row0 + row) * HID + col] = z;
                    sv.h1[(row0 + row) * HID + col] = h;
                }
            }
        }
        __syncthreads();
        // layer 2
        for (int a = 0; a < FW_MT; ++a)
            for (int r = 0; r < 16; ++r) acc[a][r] = 0.0f;
#pragma unroll
        for (int ks = 0; ks < HID / 16; ++ks) {
            bf16x8 av[FW_MT];
#pragma unroll
            for (int mt = 0; mt < FW_MT; ++mt) av[mt] = ld8(sh + (mt * 32 + li) * FW_HS + ks * 16 + g * 8);
#pragma unroll
            for (int mt = 0; mt < FW_MT; ++mt) acc[mt] = mfma32(av[mt], w2f[ks], acc[mt]);
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();   // every wavefront has read the first hidden layer
#pragma unroll
        for (int mt = 0; mt < FW_MT; ++mt) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = mt * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
                const float z = acc[mt][r] + b2c;
                const bf16_t h = f2bf_hw(silu(z));
                sh[row * FW_HS + col] = h;
                if (sv.z2 && row0 + row < n) {
                    sv.z2[(row0 + row) * HID + col] = z;
                    sv.h2[(row0 + row) * HID + col] = h;
                }
            }
        }
        __syncthreads();
        // layer 3: thread (row, quarter) sums 64 columns in order
        if (tid < 4 * FW_TM) {
            const int r = tid >> 2, q = tid & 3;
            const bf16_t* hr = sh + r * FW_HS + q * 64;
            float v = 0.0f;
            for (int c8 = 0; c8 < 64; c8 += 8) {
                const bf16x8 hv = ld8(hr + c8);
                for (int e = 0; e < 8; ++e) v += (float)hv[e] * sw3[q * 64 + c8 + e];
            }
            ypart[q][r] = v;
        }
        __syncthreads();
        if (tid < FW_TM && row0 + tid < n) {
            float v = (ypart[0][tid] + ypart[1][tid]) + (ypart[2][tid] + ypart[3][tid]) + b3;
            if (MODE == IN_REC) {
                const RecS r = ((const RecS*)in)[row0 + tid];
                v = P.out_mul * v + P.mag_mul * plog2(r.val);
                if (r.key & 1ull) v = INFINITY;
            }
            y[row0 + tid] = v;
        }
    }
}

// C[m][n] (+ split z at C + z * M * ldc) = sum_{k in split z} A[m][k] B[n][k]: BF16 operands
// with k contiguous, FP32 result.  Workgroup tile 128 x 32 (wavefront w: rows 32 w ..),
// grid (N / 32, M / 128, splits); M % 128 == 0, N % 32 == 0, K % (16 * splits) == 0.
__global__ __launch_bounds__(256) void k_gemm_nt(const bf16_t* __restrict__ A, size_t lda, const bf16_t* __restrict__ B,
                                                 size_t ldb, float* __restrict__ C, size_t ldc, int M, int K_split) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 31, g = lane >> 5;
    const size_t m0 = (size_t)blockIdx.y * 128 + w * 32, n0 = (size_t)blockIdx.x * 32;
    const size_t k0 = (size_t)blockIdx.z * K_split;
    const bf16_t* a = A + (m0 + li) * lda + k0 + g * 8;
    const bf16_t* b = B + (n0 + li) * ldb + k0 + g * 8;
    f32x16 acc;
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
#pragma unroll 4
    for (int k = 0; k < K_split; k += 16) acc = mfma32(ld8(a + k), ld8(b + k), acc);
    float* c = C + (size_t)blockIdx.z * (size_t)M * ldc;
    for (int r = 0; r < 16; ++r) c[(m0 + 8 * (r >> 2) + 4 * g + (r & 3)) * ldc + n0 + li] = acc[r];
}

// out[i] = sum_{s < S} part[s * n + i], in order of s
__global__ void k_sum_splits(const float* __restrict__ part, int S, size_t n, float* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float v = 0.0f;
        for (int s = 0; s < S; ++s) v += part[(size_t)s * n + i];
        out[i] = v;
    }
}

// BF16 transpose in[R][Cn] -> out[Cn][R] through 64 x 64 LDS tiles
__global__ void k_transpose_bf16(const bf16_t* __restrict__ in, size_t R, size_t Cn, bf16_t* __restrict__ out) {
    __shared__ bf16_t t[64][65];
    const size_t r0 = (size_t)blockIdx.y * 64, c0 = (size_t)blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int y = ty; y < 64; y += 4)
        if (r0 + y < R && c0 + tx < Cn) t[y][tx] = in[(r0 + y) * Cn + c0 + tx];
    __syncthreads();
    for (int y = ty; y < 64; y += 4)
        if (c0 + y < Cn && r0 + tx < R) out[(c0 + y) * R + r0 + tx] = t[tx][y];
}

// Column sums of a [R][HID] BF16 matrix, optionally weighted by s[r]: partial sums of 256
// consecutive rows per workgroup into part[blockIdx.x][HID] (rows in order).
__global__ __launch_bounds__(256) void k_colsum(const bf16_t* __restrict__ M, const float* __restrict__ s, size_t R,
                                                float* __restrict__ part) {
    const size_t r0 = (size_t)blockIdx.x * 256;
    const int col = threadIdx.x;
    float v = 0.0f;
    for (size_t r = r0; r < r0 + 256 && r < R; ++r) v += bf2f(M[r * HID + col]) * (s ? s[r] : 1.0f);
    part[(size_t)blockIdx.x * HID + col] = v;
}

// Output gradient of the mean squared error: dy = 2 (y - t) / n; e2[i] = (y - t)^2.
__global__ void k_out_grad(const float* __restrict__ y, const float* __restrict__ t, size_t n, float* __restrict__ dy,
                           float* __restrict__ e2) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float d = y[i] - t[i];
        dy[i] = 2.0f * d / (float)n;
        e2[i] = d * d;
    }
}

// Sum of n floats by one workgroup: thread t sums t, t + 256, ... in order, then a fixed tree.
__global__ __launch_bounds__(256) void k_sum1(const float* __restrict__ x, size_t n, float* __restrict__ out) {
    __shared__ float s[256];
    float v = 0.0f;
    for (size_t i = threadIdx.x; i < n; i += 256) v += x[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

// dz2 = bf16(dy * w3 * silu'(z2))
__global__ void k_bwd_out(const float* __restrict__ dy, const float* __restrict__ w3r, const float* __restrict__ z2,
                          size_t n, bf16_t* __restrict__ dz2) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * HID; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / HID;
        const int col = (int)(i - r * HID);
        dz2[i] = f2bf_hw(dy[r] * w3r[col] * silu_grad(z2[i]));
    }
}

// dz1 = bf16(dh1 * silu'(z1))
__global__ void k_bwd_hidden(const float* __restrict__ dh1, const float* __restrict__ z1, size_t n,
                             bf16_t* __restrict__ dz1) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * HID; i += (size_t)gridDim.x * blockDim.x)
        dz1[i] = f2bf_hw(dh1[i] * silu_grad(z1[i]));
}

// Flat gradient from the pieces: dW1 [HID][XT_ROWS] (columns < IN used), db1, dW2, db2,
// dw3, db3 (db3 from the sum of dy).
__global__ void k_grad_flat(const float* __restrict__ dw1, const float* __restrict__ db1, const float* __restrict__ dw2,
                            const float* __restrict__ db2, const float* __restrict__ dw3, const float* __restrict__ db3,
                            float* __restrict__ grad) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < MLP_NPARAM; i += gridDim.x * blockDim.x) {
        float v;
        if (i < MLP_OFF_B1) v = dw1[(i / IN) * XT_ROWS + i % IN];
        else if (i < MLP_OFF_W2) v = db1[i - MLP_OFF_B1];
        else if (i < MLP_OFF_B2) v = dw2[i - MLP_OFF_W2];
        else if (i < MLP_OFF_W3) v = db2[i - MLP_OFF_B2];
        else if (i < MLP_OFF_B3) v = dw3[i - MLP_OFF_W3];
        else v = db3[0];
        grad[i] = v;
    }
}

// Adam (bias-corrected): m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
// p -= lr * (m / c1) / (sqrt(v / c2) + eps), c1 = 1 - b1^t, c2 = 1 - b2^t.
__global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < MLP_NPARAM; i += gridDim.x * blockDim.x) {
        const float gr = grad[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gr;
        const float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// BF16 operands from the FP32 master parameters.
__global__ void k_derive(const float* __restrict__ p, bf16_t* __restrict__ w1t, bf16_t* __restrict__ w2t,
                         bf16_t* __restrict__ w2n, float* __restrict__ w3r) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < HID * HID; i += gridDim.x * blockDim.x) {
        const int n = i / HID, j = i % HID;
        const bf16_t b = f2bf(p[MLP_OFF_W2 + i]);
        w2t[i] = b;
        w2n[j * HID + n] = b;
        if (j < INP) w1t[n * INP + j] = j < IN ? f2bf(p[MLP_OFF_W1 + n * IN + j]) : (bf16_t)0;
        if (i < HID) w3r[i] = bf2f(f2bf(p[MLP_OFF_W3 + i]));
    }
}

// Features of explicit rows x[n][IN] (FP32) as BF16 [n][INP].
__global__ void k_pack_x(const float* __restrict__ x, size_t n, bf16_t* __restrict__ xb) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * INP; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / INP;
        const int c = (int)(i - r * INP);
        xb[i] = c < IN ? f2bf(x[r * IN + c]) : (bf16_t)0;
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Training batch: row i takes sample mix64(seed, step, i) mod n_samples; features as BF16
// [B][INP] and the normalised label (label - y_shift) * y_scale.
__global__ void k_batch(MlpDev P, const sfci_screen_sample_t* __restrict__ smp, size_t n_smp, int n_so, size_t B,
                        uint64_t seed, uint64_t step, float y_shift, float y_scale, bf16_t* __restrict__ xb,
                        float* __restrict__ t) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < B; i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t h = mix64(seed ^ mix64(step * 0x100000001B3ull + i));
        const sfci_screen_sample_t s = smp[h % n_smp];
        for (int c = 0; c < INP; ++c) xb[i * INP + c] = f2bf(feature(c, s.det, n_so, s.cls, s.logv, P));
        t[i] = (s.label - y_shift) * y_scale;
    }
}

// Features of samples as FP32 [n][IN] (inspection and tests).
__global__ void k_sample_features(MlpDev P, const sfci_screen_sample_t* __restrict__ smp, size_t n, int n_so,
                                  float* __restrict__ x) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * IN; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / IN;
        const int c = (int)(i - r * IN);
        const sfci_screen_sample_t s = smp[r];
        x[i] = feature(c, s.det, n_so, s.cls, s.logv, P);
    }
}

}  // namespace sfci

struct sfci_mlp {
    Ctx* c = nullptr;
    DevBuf p, m, v, grad;
    DevBuf w1t, w2t, w2n, w3r;
    float x_shift = 0.0f, x_scale = 1.0f, y_shift = 0.0f, y_scale = 1.0f;
    int64_t adam_t = 0;
    // training scratch (sized for the batch)
    size_t cap_rows = 0;
    DevBuf xb, xt, t, y, dy, e2, z1, h1, z2, h2, h1t, dz2, dz2t, dz1, dz1t, dh1, gpart, dw1, dw2, db1, db2, dw3, db3,
        loss, smp;

    MlpDev dev() const {
        MlpDev d;
        d.w1t = (const bf16_t*)w1t.p;
        d.b1 = (const float*)p.p + MLP_OFF_B1;
        d.w2t = (const bf16_t*)w2t.p;
        d.b2 = (const float*)p.p + MLP_OFF_B2;
        d.w3r = (const float*)w3r.p;
        d.b3 = (const float*)p.p + MLP_OFF_B3;
        d.x_shift = x_shift;
        d.x_scale = x_scale;
        d.out_mul = 1.0f;
        d.mag_mul = 0.0f;
        return d;
    }
    void derive() {
        k_derive<<<256, 256, 0, c->stream>>>((const float*)p.p, (bf16_t*)w1t.p, (bf16_t*)w2t.p, (bf16_t*)w2n.p,
                                             (float*)w3r.p);
        HIPCHK(hipGetLastError());
    }
};

namespace {

int grid1(size_t n, int per = 256, int max_blocks = 4096) {
    size_t b = (n + per - 1) / per;
    return (int)std::max<size_t>(1, std::min<size_t>(b, (size_t)max_blocks));
}

void fwd_launch(sfci_mlp* net, int mode, const void* in, size_t n, float* y, FwdSave sv, bool with_magnitude) {
    Ctx* c = net->c;
    if (n == 0) return;
    const size_t tiles = (n + FW_TM - 1) / FW_TM;
    const int grid = (int)std::min<size_t>(tiles, 256);   // one resident workgroup per CU
    MlpDev P = net->dev();
    if (with_magnitude) {
        // predicted ln|y| of the target plus ln|v| of the candidate
        P.out_mul = 1.0f / net->y_scale;
        P.mag_mul = 0.69314718f;
    }
    if (mode == IN_X)
        k_mlp_fwd<IN_X><<<grid, FW_THREADS, 0, c->stream>>>(P, c->n_so, in, n, y, sv);
    else if (mode == IN_REC)
        k_mlp_fwd<IN_REC><<<grid, FW_THREADS, 0, c->stream>>>(P, c->n_so, in, n, y, sv);
    else
        k_mlp_fwd<IN_SAMPLE><<<grid, FW_THREADS, 0, c->stream>>>(P, c->n_so, in, n, y, sv);
    HIPCHK(hipGetLastError());
}

// Training scratch for batches of up to `rows` rows (rows % 256 == 0).
void ensure_rows(sfci_mlp* net, size_t rows) {
    if (rows <= net->cap_rows) return;
    net->xb.get<bf16_t>(rows * INP);
    bf16_t* xt = net->xt.get<bf16_t>(XT_ROWS * rows);
    HIPCHK(hipMemsetAsync(xt, 0, XT_ROWS * rows * sizeof(bf16_t), net->c->stream));
    net->t.get<float>(rows);
    net->y.get<float>(rows);
    net->dy.get<float>(rows);
    net->e2.get<float>(rows);
    net->z1.get<float>(rows * HID);
    net->h1.get<bf16_t>(rows * HID);
    net->z2.get<float>(rows * HID);
    net->h2.get<bf16_t>(rows * HID);
    net->h1t.get<bf16_t>(rows * HID);
    net->dz2.get<bf16_t>(rows * HID);
    net->dz2t.get<bf16_t>(rows * HID);
    net->dz1.get<bf16_t>(rows * HID);
    net->dz1t.get<bf16_t>(rows * HID);
    net->dh1.get<float>(rows * HID);
    // split-K partials of the weight gradients and the column-sum partials
    net->gpart.get<float>(std::max<size_t>((size_t)64 * HID * HID, rows));
    net->cap_rows = rows;
}

int splits_for(size_t K) {
    int s = 1;
    while (s < 64 && K / (size_t)(2 * s) >= 1024 && (K % (size_t)(32 * s)) == 0) s *= 2;
    return s;
}

// C[M][N] = A[M][K] B[N][K]^T (split-K with an ordered sum of the partials)
void gemm_nt(sfci_mlp* net, const bf16_t* A, size_t lda, const bf16_t* B, size_t ldb, float* C, size_t M, size_t N,
             size_t K) {
    Ctx* c = net->c;
    if (M % 128 || N % 32 || K % 16) throw HipError("gemm_nt: shape not tiled", -1);
    const int S = splits_for(K);
    float* dst = C;
    if (S > 1) dst = net->gpart.get<float>((size_t)S * M * N);
    dim3 grid((unsigned)(N / 32), (unsigned)(M / 128), (unsigned)S);
    k_gemm_nt<<<grid, 256, 0, c->stream>>>(A, lda, B, ldb, dst, N, (int)M, (int)(K / S));
    HIPCHK(hipGetLastError());
    if (S > 1) {
        k_sum_splits<<<grid1(M * N), 256, 0, c->stream>>>(dst, S, M * N, C);
        HIPCHK(hipGetLastError());
    }
}

void colsum(sfci_mlp* net, const bf16_t* Mx, const float* s, size_t R, float* out) {
    Ctx* c = net->c;
    const int nb = (int)((R + 255) / 256);
    float* part = net->gpart.get<float>((size_t)nb * HID);
    k_colsum<<<nb, 256, 0, c->stream>>>(Mx, s, R, part);
    k_sum_splits<<<1, 256, 0, c->stream>>>(part, nb, HID, out);
    HIPCHK(hipGetLastError());
}

// Forward + backward over the rows already in net->xb / net->t (n rows, n % 256 == 0);
// leaves the flat gradient in net->grad and the mean squared error in net->loss.
void forward_backward(sfci_mlp* net, size_t n) {
    Ctx* c = net->c;
    const hipStream_t st = c->stream;
    FwdSave sv{(float*)net->z1.p, (bf16_t*)net->h1.p, (float*)net->z2.p, (bf16_t*)net->h2.p};
    fwd_launch(net, IN_X, net->xb.p, n, (float*)net->y.p, sv, false);
    k_out_grad<<<grid1(n), 256, 0, st>>>((const float*)net->y.p, (const float*)net->t.p, n, (float*)net->dy.p,
                                         (float*)net->e2.p);
    float* loss = net->loss.get<float>(2);
    k_sum1<<<1, 256, 0, st>>>((const float*)net->e2.p, n, loss);
    k_sum1<<<1, 256, 0, st>>>((const float*)net->dy.p, n, net->db3.get<float>(1));
    // output layer
    k_bwd_out<<<grid1(n * HID), 256, 0, st>>>((const float*)net->dy.p, (const float*)net->w3r.p,
                                              (const float*)net->z2.p, n, (bf16_t*)net->dz2.p);
    colsum(net, (const bf16_t*)net->h2.p, (const float*)net->dy.p, n, net->dw3.get<float>(HID));
    colsum(net, (const bf16_t*)net->dz2.p, nullptr, n, net->db2.get<float>(HID));
    // hidden layer 2 -> 1
    gemm_nt(net, (const bf16_t*)net->dz2.p, HID, (const bf16_t*)net->w2n.p, HID, (float*)net->dh1.p, n, HID, HID);
    k_bwd_hidden<<<grid1(n * HID), 256, 0, st>>>((const float*)net->dh1.p, (const float*)net->z1.p, n,
                                                 (bf16_t*)net->dz1.p);
    colsum(net, (const bf16_t*)net->dz1.p, nullptr, n, net->db1.get<float>(HID));
    // weight gradients: transposed operands (batch contiguous)
    dim3 tg((unsigned)(HID / 64), (unsigned)((n + 63) / 64));
    k_transpose_bf16<<<tg, 256, 0, st>>>((const bf16_t*)net->dz2.p, n, HID, (bf16_t*)net->dz2t.p);
    k_transpose_bf16<<<tg, 256, 0, st>>>((const bf16_t*)net->h1.p, n, HID, (bf16_t*)net->h1t.p);
    k_transpose_bf16<<<tg, 256, 0, st>>>((const bf16_t*)net->dz1.p, n, HID, (bf16_t*)net->dz1t.p);
    dim3 tx((unsigned)((INP + 63) / 64), (unsigned)((n + 63) / 64));
    k_transpose_bf16<<<tx, 256, 0, st>>>((const bf16_t*)net->xb.p, n, INP, (bf16_t*)net->xt.p);
    HIPCHK(hipGetLastError());
    gemm_nt(net, (const bf16_t*)net->dz2t.p, n, (const bf16_t*)net->h1t.p, n, net->dw2.get<float>(HID * HID), HID,
            HID, n);
    gemm_nt(net, (const bf16_t*)net->dz1t.p, n, (const bf16_t*)net->xt.p, n, net->dw1.get<float>(HID * XT_ROWS), HID,
            XT_ROWS, n);
    k_grad_flat<<<grid1(MLP_NPARAM), 256, 0, st>>>((const float*)net->dw1.p, (const float*)net->db1.p,
                                                   (const float*)net->dw2.p, (const float*)net->db2.p,
                                                   (const float*)net->dw3.p, (const float*)net->db3.p,
                                                   net->grad.get<float>(MLP_NPARAM));
    HIPCHK(hipGetLastError());
}

void adam_step(sfci_mlp* net, float lr, float b1, float b2, float eps) {
    net->adam_t += 1;
    const float c1 = 1.0f - powf(b1, (float)net->adam_t), c2 = 1.0f - powf(b2, (float)net->adam_t);
    k_adam<<<grid1(MLP_NPARAM), 256, 0, net->c->stream>>>((float*)net->p.p, (float*)net->m.p, (float*)net->v.p,
                                                          (const float*)net->grad.p, lr, b1, b2, eps, c1, c2);
    HIPCHK(hipGetLastError());
    net->derive();
}

// rows of explicit features: x[n][IN] FP32 -> BF16 [n_pad][INP] in net->xb (zero rows past n)
size_t load_x(sfci_mlp* net, const float* x, size_t n) {
    Ctx* c = net->c;
    const size_t n_pad = (n + 255) / 256 * 256;
    ensure_rows(net, n_pad);
    DevBuf tmp;
    float* dx = tmp.get<float>(n_pad * IN);
    HIPCHK(hipMemsetAsync(dx, 0, n_pad * IN * sizeof(float), c->stream));
    HIPCHK(hipMemcpyAsync(dx, x, n * IN * sizeof(float), hipMemcpyHostToDevice, c->stream));
    k_pack_x<<<grid1(n_pad * INP), 256, 0, c->stream>>>(dx, n_pad, (bf16_t*)net->xb.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    tmp.release();
    return n_pad;
}

}  // namespace

namespace sfci {
void mlp_score_records(Ctx* c, const sfci_mlp* net, const RecS* rec, size_t n, float* score, bool with_magnitude) {
    int tk = kt_begin(c, K_MLP_SCORE, 20.0 * n);
    fwd_launch(const_cast<sfci_mlp*>(net), IN_REC, rec, n, score, FwdSave{nullptr, nullptr, nullptr, nullptr},
               with_magnitude);
    kt_end(c, tk);
}
}  // namespace sfci

extern "C" {

int sfci_mlp_create(sfci_ctx* X, uint64_t seed, sfci_mlp** out) {
    if (!X || !out) return SFCI_E_ARG;
    *out = nullptr;
    sfci_mlp* net = new sfci_mlp();
    net->c = &X->c;
    int rc = guarded(&X->c, [&] {
        // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases, from a 64-bit LCG
        std::vector<float> p(MLP_NPARAM, 0.0f);
        uint64_t s = seed * 6364136223846793005ull + 1442695040888963407ull;
        auto uni = [&]() {
            s = s * 6364136223846793005ull + 1442695040888963407ull;
            return (double)(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
        };
        for (int i = 0; i < HID * IN; ++i) p[MLP_OFF_W1 + i] = (float)(uni() / std::sqrt((double)IN));
        for (int i = 0; i < HID * HID; ++i) p[MLP_OFF_W2 + i] = (float)(uni() / std::sqrt((double)HID));
        for (int i = 0; i < HID; ++i) p[MLP_OFF_W3 + i] = (float)(uni() / std::sqrt((double)HID));
        float* dp = net->p.get<float>(MLP_NPARAM);
        HIPCHK(hipMemcpyAsync(dp, p.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, X->c.stream));
        HIPCHK(hipMemsetAsync(net->m.get<float>(MLP_NPARAM), 0, MLP_NPARAM * sizeof(float), X->c.stream));
        HIPCHK(hipMemsetAsync(net->v.get<float>(MLP_NPARAM), 0, MLP_NPARAM * sizeof(float), X->c.stream));
        net->grad.get<float>(MLP_NPARAM);
        net->w1t.get<bf16_t>(HID * INP);
        net->w2t.get<bf16_t>(HID * HID);
        net->w2n.get<bf16_t>(HID * HID);
        net->w3r.get<float>(HID);
        net->derive();
        HIPCHK(hipStreamSynchronize(X->c.stream));
    });
    if (rc) {
        delete net;
        return rc;
    }
    *out = net;
    return SFCI_OK;
}

void sfci_mlp_destroy(sfci_mlp* net) {
    if (!net) return;
    (void)hipSetDevice(net->c->device);
    (void)hipStreamSynchronize(net->c->stream);
    DevBuf* bufs[] = {&net->p,  &net->m,   &net->v,   &net->grad, &net->w1t,  &net->w2t, &net->w2n,  &net->w3r,
                      &net->xb, &net->xt,  &net->t,   &net->y,    &net->dy,   &net->e2,  &net->z1,   &net->h1,
                      &net->z2, &net->h2,  &net->h1t, &net->dz2,  &net->dz2t, &net->dz1, &net->dz1t, &net->dh1,
                      &net->gpart, &net->dw1, &net->dw2, &net->db1, &net->db2, &net->dw3, &net->db3, &net->loss,
                      &net->smp};
    for (DevBuf* b : bufs) b->release();
    delete net;
}

int sfci_mlp_set_params(sfci_mlp* net, const float* params) {
    if (!net || !params) return SFCI_E_ARG;
    return guarded(net->c, [&] {
        HIPCHK(hipMemcpyAsync(net->p.p, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, net->c->stream));
        HIPCHK(hipMemsetAsync(net->m.p, 0, MLP_NPARAM * sizeof(float), net->c->stream));
        HIPCHK(hipMemsetAsync(net->v.p, 0, MLP_NPARAM * sizeof(float), net->c->stream));
        net->adam_t = 0;
        net->derive();
        HIPCHK(hipStreamSynchronize(net->c->stream));
    });
}

int sfci_mlp_get_params(const sfci_mlp* net, float* params) {
    if (!net || !params) return SFCI_E_ARG;
    return guarded(net->c, [&] {
        HIPCHK(hipMemcpyAsync(params, net->p.p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, net->c->stream));
        HIPCHK(hipStreamSynchronize(net->c->stream));
    });
}

int sfci_mlp_set_norm(sfci_mlp* net, float x_shift, float x_scale, float y_shift, float y_scale) {
    if (!net || !(x_scale > 0.0f) || !(y_scale > 0.0f)) return SFCI_E_ARG;
    net->x_shift = x_shift;
    net->x_scale = x_scale;
    net->y_shift = y_shift;
    net->y_scale = y_scale;
    return SFCI_OK;
}

int sfci_mlp_forward(sfci_mlp* net, const float* x, int64_t n, float* y) {
    if (!net || n < 0 || (n > 0 && (!x || !y))) return SFCI_E_ARG;
    if (n == 0) return SFCI_OK;
    return guarded(net->c, [&] {
        const size_t np = load_x(net, x, (size_t)n);
        fwd_launch(net, IN_X, net->xb.p, np, (float*)net->y.p, FwdSave{nullptr, nullptr, nullptr, nullptr}, false);
        HIPCHK(hipMemcpyAsync(y, net->y.p, n * sizeof(float), hipMemcpyDeviceToHost, net->c->stream));
        HIPCHK(hipStreamSynchronize(net->c->stream));
    });
}

int sfci_mlp_gradient(sfci_mlp* net, const float* x, const float* t, int64_t n, float* grad, float* loss) {
    if (!net || !x || !t || !grad || n < 1 || n % 256) return SFCI_E_ARG;
    return guarded(net->c, [&] {
        load_x(net, x, (size_t)n);
        HIPCHK(hipMemcpyAsync(net->t.p, t, n * sizeof(float), hipMemcpyHostToDevice, net->c->stream));
        forward_backward(net, (size_t)n);
        HIPCHK(hipMemcpyAsync(grad, net->grad.p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, net->c->stream));
        float l = 0.0f;
        HIPCHK(hipMemcpyAsync(&l, net->loss.p, sizeof(float), hipMemcpyDeviceToHost, net->c->stream));
        HIPCHK(hipStreamSynchronize(net->c->stream));
        if (loss) *loss = l / (float)n;
    });
}

int sfci_mlp_adam_step(sfci_mlp* net, const float* grad, float lr, float beta1, float beta2, float eps) {
    if (!net || !grad) return SFCI_E_ARG;
    return guarded(net->c, [&] {
        HIPCHK(hipMemcpyAsync(net->grad.p, grad, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, net->c->stream));
        adam_step(net, lr, beta1, beta2, eps);
        HIPCHK(hipStreamSynchronize(net->c->stream));
    });
}

int sfci_mlp_train(sfci_mlp* net, const sfci_screen_sample_t* samples, int64_t n, int steps, int batch, float lr,
                   uint64_t seed, float* loss_curve) {
    if (!net || !samples || n < 1 || steps < 0 || batch < 256 || batch % 256) return SFCI_E_ARG;
    Ctx* c = net->c;
    return guarded(c, [&] {
        sfci_screen_sample_t* ds = net->smp.get<sfci_screen_sample_t>((size_t)n);
        HIPCHK(hipMemcpyAsync(ds, samples, n * sizeof(sfci_screen_sample_t), hipMemcpyHostToDevice, c->stream));
        ensure_rows(net, (size_t)batch);
        float* curve = nullptr;
        DevBuf lc;
        if (loss_curve && steps > 0) curve = lc.get<float>((size_t)steps);
        const MlpDev P = net->dev();
        for (int s = 0; s < steps; ++s) {
            int tk = kt_begin(c, K_MLP_TRAIN, 0.0);
            k_batch<<<grid1((size_t)batch), 256, 0, c->stream>>>(P, ds, (size_t)n, c->n_so, (size_t)batch, seed,
                                                                 (uint64_t)s, net->y_shift, net->y_scale,
                                                                 (bf16_t*)net->xb.p, (float*)net->t.p);
            forward_backward(net, (size_t)batch);
            adam_step(net, lr, 0.9f, 0.999f, 1e-8f);
            kt_end(c, tk);
            if (curve) HIPCHK(hipMemcpyAsync(curve + s, net->loss.p, sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        }
        if (curve) {
            std::vector<float> h((size_t)steps);
            HIPCHK(hipMemcpyAsync(h.data(), curve, steps * sizeof(float), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int s = 0; s < steps; ++s) loss_curve[s] = h[s] / (float)batch;
        }
        kt_resolve(c);
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int sfci_mlp_predict_samples(sfci_mlp* net, const sfci_screen_sample_t* samples, int64_t n, float* y) {
    if (!net || n < 0 || (n > 0 && (!samples || !y))) return SFCI_E_ARG;
    if (n == 0) return SFCI_OK;
    Ctx* c = net->c;
    return guarded(c, [&] {
        sfci_screen_sample_t* ds = net->smp.get<sfci_screen_sample_t>((size_t)n);
        HIPCHK(hipMemcpyAsync(ds, samples, n * sizeof(sfci_screen_sample_t), hipMemcpyHostToDevice, c->stream));
        DevBuf out;
        float* dy = out.get<float>((size_t)n);
        fwd_launch(net, IN_SAMPLE, ds, (size_t)n, dy, FwdSave{nullptr, nullptr, nullptr, nullptr}, false);
        HIPCHK(hipMemcpyAsync(y, dy, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        out.release();
    });
}

int sfci_mlp_sample_features(sfci_mlp* net, const sfci_screen_sample_t* samples, int64_t n, float* x) {
    if (!net || n < 0 || (n > 0 && (!samples || !x))) return SFCI_E_ARG;
    if (n == 0) return SFCI_OK;
    Ctx* c = net->c;
    return guarded(c, [&] {
        sfci_screen_sample_t* ds = net->smp.get<sfci_screen_sample_t>((size_t)n);
        HIPCHK(hipMemcpyAsync(ds, samples, n * sizeof(sfci_screen_sample_t), hipMemcpyHostToDevice, c->stream));
        DevBuf out;
        float* dx = out.get<float>((size_t)n * IN);
        k_sample_features<<<grid1((size_t)n * IN), 256, 0, c->stream>>>(net->dev(), ds, (size_t)n, c->n_so, dx);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(x, dx, n * IN * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        out.release();
    });
}

}  // extern "C"
