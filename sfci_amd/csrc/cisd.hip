// CISD start vector: calc_CISD (src/sparse_solver.cpp:15-114) on the GPU.
//
// The block is ref_det plus every determinant its H.v column reaches with a nonzero
// element (the reference collects the targets of the integral-list entries that
// OffDiagGen_bit_e emits for ref_det, src/bitstr.cpp:304-431; the two sets differ only
// where nonzero list entries cancel exactly).  Its dense Hamiltonian is filled from the
// same unique-emission generator the H.v uses (one record per connected determinant,
// exact element), then diagonalized by rocSOLVER's dsyevd (divide and conquer) in place
// of Eigen's SelfAdjointEigenSolver.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <vector>

#include "internal.h"

namespace sfci {
namespace {

__global__ void k_block_flags(const uint64_t* __restrict__ k, const double* __restrict__ v, size_t n, uint64_t ref,
                              uint32_t* __restrict__ fl) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        fl[i] = (k[i] == ref || v[i] != 0.0) ? 1u : 0u;
}

__global__ void k_block_compact(const uint64_t* __restrict__ k, size_t n, const uint32_t* __restrict__ fl,
                                const uint64_t* __restrict__ pos, uint64_t* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        if (fl[i]) out[pos[i]] = k[i];
}

// dense key (colex rank of the spin-0 string) * dim + (rank of the spin-1 string), and the
// row index as the value of a (key, value) sort
__global__ void k_block_dense(HamDev H, const uint64_t* __restrict__ dets, size_t n, uint64_t* __restrict__ dk,
                              double* __restrict__ idx) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint64_t r0 = colex_rank(spin_string(dets[i], 0), H.binom);
        uint64_t r1 = colex_rank(spin_string(dets[i], 1), H.binom);
        dk[i] = r0 * H.dim + r1;
        idx[i] = (double)i;
    }
}

// records of source row i (at i * n_conn + e) whose target lies in the block -> A[i][p]
__global__ void k_block_fill(const Rec64* __restrict__ rec, uint64_t n_conn, size_t n,
                             const uint64_t* __restrict__ sorted_dense, const double* __restrict__ sorted_idx,
                             const double* __restrict__ dterm, double* __restrict__ A) {
    const size_t total = n * n_conn;
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < total; r += (size_t)gridDim.x * blockDim.x) {
        const size_t i = r / n_conn;
        const Rec64 x = rec[r];
        const uint64_t key = x.key;
        if (key & 1) {
            A[i * n + i] = dterm[i];
            continue;
        }
        const uint64_t t = key >> 1;
        size_t lo = 0, hi = n;
        while (lo < hi) {
            size_t mid = (lo + hi) >> 1;
            if (sorted_dense[mid] < t) lo = mid + 1;
            else hi = mid;
        }
        if (lo < n && sorted_dense[lo] == t) A[i * n + (size_t)sorted_idx[lo]] = x.val;
    }
}

int grid1(size_t n) {
    size_t g = (n + SFCI_BLOCK - 1) / SFCI_BLOCK;
    return (int)std::max<size_t>(1, std::min<size_t>(g, 256 * 16));
}

struct RocblasHandle {
    rocblas_handle h = nullptr;
    ~RocblasHandle() {
        if (h) rocblas_destroy_handle(h);
    }
};

}  // namespace

size_t dev_cisd(Ctx* c, uint64_t ref_det, int state_idx, size_t out_cap, uint64_t* ok, double* ov,
                double* eigenvalue) {
    const HamDev& H = c->ham;
    // 1. the block: ref_det and every target of its H.v column with a nonzero element
    uint64_t* rk = c->sel_a.get<uint64_t>(1);
    HIPCHK(hipMemcpyAsync(rk, &ref_det, 8, hipMemcpyHostToDevice, c->stream));
    double* rv1 = c->lin_v[1].get<double>(1);
    double one = 1.0, d = 0.0;
    HIPCHK(hipMemcpyAsync(rv1, &one, 8, hipMemcpyHostToDevice, c->stream));
    HvAll col;
    dev_hv_all(c, rk, rv1, 1, 0.0, 1.0, &d, &col);
    const size_t nc = col.n;
    uint32_t* fl = c->sel_b.get<uint32_t>(nc);
    uint64_t* pos = c->sel_c.get<uint64_t>(nc + 1);
    k_block_flags<<<grid1(nc), SFCI_BLOCK, 0, c->stream>>>(col.k, col.v, nc, ref_det, fl);
    exclusive_scan_u32(c, fl, nc, pos, pos + nc);
    uint64_t n64 = 0;
    HIPCHK(hipMemcpyAsync(&n64, pos + nc, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)n64;
    if (n == 0 || (size_t)state_idx >= n) throw HipError("cisd: state index outside the CISD block", -1);
    if (n > out_cap) throw HipError("cisd: output vector smaller than the CISD block", -4);
    if (n > (size_t)0x7fffffff) throw HipError("cisd: block too large for the dense eigensolver", -4);
    uint64_t* dets = c->lin_k[0].get<uint64_t>(n);
    k_block_compact<<<grid1(nc), SFCI_BLOCK, 0, c->stream>>>(col.k, nc, fl, pos, dets);
    // 2. dense keys of the block, sorted, with their row index
    uint64_t* dk = c->tmp_k[0].get<uint64_t>(n);
    double* di = c->tmp_v[0].get<double>(n);
    uint64_t* dk2 = c->lin_k[1].get<uint64_t>(n);
    double* di2 = c->lin_v[1].get<double>(n);
    k_block_dense<<<grid1(n), SFCI_BLOCK, 0, c->stream>>>(H, dets, n, dk, di);
    uint64_t* sk = dk;
    double* sv = di;
    radix_sort_pairs(c, &sk, &sv, dk2, di2, n, bits_for(H.dim * H.dim));
    // 3. every connection of every block determinant; entries inside the block fill A
    const Rec64* recs;
    const double* dterm;
    dev_connection_records(c, dets, n, &recs, &dterm);
    DevBuf A, W;
    double* a = A.get<double>(n * n);
    double* evals = W.get<double>(2 * n + 1);
    HIPCHK(hipMemsetAsync(a, 0, n * n * sizeof(double), c->stream));
    k_block_fill<<<grid1(n * H.n_conn), SFCI_BLOCK, 0, c->stream>>>(recs, H.n_conn, n, sk, sv, dterm, a);
    HIPCHK(hipGetLastError());
    // 4. all eigenpairs (ascending); column state_idx of the column-major result
    RocblasHandle hdl;
    if (rocblas_create_handle(&hdl.h) != rocblas_status_success) throw HipError("cisd: rocblas_create_handle", -2);
    rocblas_set_stream(hdl.h, c->stream);
    rocblas_int* info = (rocblas_int*)(evals + 2 * n);
    if (rocsolver_dsyevd(hdl.h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n, a, (rocblas_int)n, evals,
                         evals + n, info) != rocblas_status_success)
        throw HipError("cisd: rocsolver_dsyevd failed", -2);
    rocblas_int info_h = 0;
    std::vector<double> vec(n);
    HIPCHK(hipMemcpyAsync(&info_h, info, sizeof(info_h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(eigenvalue, evals + state_idx, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(vec.data(), a + (size_t)state_idx * n, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (info_h != 0) throw HipError("cisd: dsyevd did not converge", -2);
    // sign: the largest |c| (first, i.e. smallest key, among ties) positive
    size_t imax = 0;
    for (size_t i = 1; i < n; ++i)
        if (std::fabs(vec[i]) > std::fabs(vec[imax])) imax = i;
    if (vec[imax] < 0)
        for (double& x : vec) x = -x;
    HIPCHK(hipMemcpyAsync(ok, dets, n * 8, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(ov, vec.data(), n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return n;
}

}  // namespace sfci
