/*
 * sfci_gpu.h -- C-ABI of the MI355X (gfx950) SFCI H.v library, libsfci_hip.so.
 *
 * The reference (wlj89/SFCI) is one C++ binary with no plugin or FFI layer; its hot
 * path is the in-process call
 *     void SpMVMtply_bucket(dynVec<u64,double>& in, dynVec<u64,double>& out,
 *                           double lambda, double coeff, double& diag);
 * (declared src/all.h:642, defined src/functions.cpp:511-804) made by the Lanczos
 * driver (src/sparse_solver.cpp:306, :358), plus the dynVec methods the driver uses
 * (src/all.h:150-625) and calc_energy (src/functions.cpp:1272-1313).  This header is
 * the drop-in for that call boundary: plain pointers and sizes, device-resident sparse
 * vectors, integer return codes.
 *
 * Conventions
 *   - Every function returns SFCI_OK (0) or a negative SFCI_E_* code; the message of
 *     the last failure is available from sfci_last_error().  No C++ exception crosses
 *     the ABI.
 *   - One context = one GPU = one HIP stream.  A context is not thread-safe.
 *   - Sparse vectors hold (key, coefficient) pairs sorted by ascending key, like
 *     dynVec; keys are determinant bit strings (bit p = spin orbital p, the numbering
 *     of the reference's _IN files after ReadElements' "-2", src/functions.cpp:329-352).
 *   - Every |c| truncation keeps the largest |c|; among equal |c| the smaller key.
 *   - Sums are exact (128-bit fixed point, unit 2^-84) and rounded once, so results do
 *     not depend on thread counts or summation order.
 */
#ifndef SFCI_GPU_H
#define SFCI_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFCI_OK 0
#define SFCI_E_ARG (-1)      /* invalid argument, malformed input file or vector */
#define SFCI_E_HIP (-2)      /* HIP runtime error (allocation, launch) */
#define SFCI_E_OVERFLOW (-3) /* device staging overflow that regrowing did not fix */
#define SFCI_E_CAPACITY (-4) /* problem does not fit the workspace limit */
#define SFCI_E_RANGE (-5)    /* a term left the fixed-point range (|x| >= 2^42) */
#define SFCI_E_IO (-6)       /* integral file could not be opened */

typedef struct sfci_ctx sfci_ctx;
typedef struct sfci_vec sfci_vec;

typedef struct {
    int n_so;             /* spin orbitals (num_orb) */
    int n_elec;           /* electrons (num_e), half of each spin */
    uint64_t string_dim;  /* C(n_so/2, n_elec/2): the reference's blockDim (src/main.cpp:183) */
    uint64_t n_conn;      /* connections per determinant including the diagonal (U + 1) */
    double e_core;        /* constant block of the _IN file */
} sfci_info_t;

typedef struct {
    uint64_t n_src;         /* input determinants used (top-Nd of the input) */
    uint64_t n_candidates;  /* connections generated = n_src * n_conn */
    uint64_t n_unique;      /* distinct target determinants before the cut */
    uint64_t n_out;         /* output length */
    uint64_t n_passes;      /* target-range passes */
    double ms_generate, ms_sort, ms_reduce, ms_select, ms_total;
    uint64_t n_fallback;    /* passes merged by the full sort although planned for hash buckets (a
                               partition region or a 256-class bucket overflowed) */
    int hash_bits;          /* hash bits of the last pass's bucket merge (0: full sort) */
    uint64_t n_kept;        /* screened H.v: candidates kept for the merge (n_candidates = emitted) */
    double screen_eps;      /* screened H.v: prefilter used (raised when the list would not fit) */
    double ms_score;        /* screened H.v: scoring time */
    uint64_t n_split;       /* bucket merge: hash classes split because their LDS table filled
                               (each split rereads the bucket's records once per subclass) */
    uint64_t n_str0, n_str1; /* distinct spin-0 / spin-1 source strings whose same-spin double
                               tables the generator read (0: it enumerated those doubles itself) */
} sfci_hv_stats_t;

/* Context from a reference _IN integral file: replaces load_integrals/ReadElements
 * (src/functions.cpp:297-418) and the parameter block of main (src/main.cpp:183-223).
 * n_so <= 64, n_elec even.  device: HIP device ordinal. */
int sfci_create(const char* in_path, int n_so, int n_elec, int device, sfci_ctx** out);

/* Same from entries already in memory, with the _IN numbering (coeff, m, n, p, q; the
 * constant as 0 0 0 0, one-body as m n 0 0): idx holds 4 int32 per entry. */
int sfci_create_from_entries(int n_so, int n_elec, int64_t n_entries, const double* coeff, const int32_t* idx,
                             int device, sfci_ctx** out);

void sfci_destroy(sfci_ctx* ctx);
const char* sfci_last_error(const sfci_ctx* ctx);
int sfci_info(const sfci_ctx* ctx, sfci_info_t* info);

/* |y| > cut filter of H.v output; default 1e-5 (SpMSpVFinalResCut, src/main.cpp:196). */
int sfci_set_cut(sfci_ctx* ctx, double cut);
/* Candidate precision of H.v (no reference counterpart; the reference is FP64 only,
 * src/all.h:76).  SFCI_PREC_FP64: every product in double.  SFCI_PREC_MIXED: each
 * off-diagonal H_ij * coeff * c_j is rounded to float (product of the float-rounded
 * factors) and moved as 12-byte records; sums stay exact, the diagonal term, the output
 * vector and all Lanczos recurrences stay FP64. */
#define SFCI_PREC_FP64 0
#define SFCI_PREC_MIXED 1
int sfci_set_precision(sfci_ctx* ctx, int precision);

/* Device bytes the H.v candidate buffers of one target pass may use. */
int sfci_set_workspace(sfci_ctx* ctx, uint64_t bytes);
int sfci_hv_stats(const sfci_ctx* ctx, sfci_hv_stats_t* st);
/* How H.v groups candidates by target (no reference counterpart; a tuning and test knob):
 * mode 0 = automatic (hash buckets for large passes), 1 = full radix sort only,
 * 2 = hash buckets with exactly hash_bits bits (1..16), 3 = as 2 with each LDS table counted
 * full at 230 occupied slots (1/16 of a full-key table's usual occupancy; forces nested class
 * splits), 4 = as 2 with full-key
 * records (no compact 12/8-byte records).  Results are identical in every mode. */
int sfci_set_merge_mode(sfci_ctx* ctx, int mode, int hash_bits);
/* How H.v evaluates same-spin doubles (no reference counterpart; a tuning and test knob):
 * 0 = automatic, 1 = every source enumerates and evaluates its own (OffDiagGen_bit's way,
 * src/bitstr.cpp:236-300), 2 = from per-string tables always: each distinct source spin
 * string's doubles are evaluated once per call and streamed to every source with that
 * string.  Results are identical in every mode. */
int sfci_set_gen_tables(sfci_ctx* ctx, int mode);
int sfci_sync(sfci_ctx* ctx);

/* Per-kernel HIP-event timings accumulated on the context's stream since the last reset
 * (no reference counterpart: the reference only prints wall-clock timers,
 * src/functions.cpp:800-801).  Kernel ids 0..SFCI_KERNEL_COUNT-1 are named by
 * sfci_kernel_name(); bytes are the algorithmic bytes each launch reads and writes. */
#define SFCI_KERNEL_COUNT 18
typedef struct {
    double ms;
    uint64_t launches;
    double bytes;
} sfci_kernel_stat_t;
int sfci_kernel_stats(sfci_ctx* ctx, sfci_kernel_stat_t* out, int n, int reset);
const char* sfci_kernel_name(int id);

/* Device sparse vector with fixed capacity: dynVec(cap) (src/all.h:170-179). */
int sfci_vec_create(sfci_ctx* ctx, int64_t cap, sfci_vec** out);
void sfci_vec_destroy(sfci_vec* v);
int64_t sfci_vec_len(const sfci_vec* v);
int64_t sfci_vec_cap(const sfci_vec* v);
/* Host -> device; keys strictly ascending, n <= cap. */
int sfci_vec_upload(sfci_vec* v, const uint64_t* keys, const double* vals, int64_t n);
/* Device -> host; copies min(len, max_n) entries. */
int sfci_vec_download(const sfci_vec* v, uint64_t* keys, double* vals, int64_t max_n, int64_t* n_out);
/* Device pointers of the key and value arrays (for zero-copy interop). */
int sfci_vec_device_ptrs(sfci_vec* v, uint64_t** keys, double** vals);
/* Declare the length after writing through the device pointers (keys must be sorted). */
int sfci_vec_set_len(sfci_vec* v, int64_t len);
/* dynVec::naiveCopy (src/all.h:610-623): dst <- src, src.len <= dst.cap. */
int sfci_vec_copy(sfci_vec* dst, const sfci_vec* src);
/* dynVec::setZeroVec (src/all.h:244-247). */
int sfci_vec_clear(sfci_vec* v);

/* SpMVMtply_bucket (src/functions.cpp:511-804):
 *   kept = top-nd of in by |c|
 *   out  = top-{out.cap} of { t : |y_t| > cut }, y = coeff * (H - lambda) * kept, key-sorted
 *   diag = <kept| coeff (H - lambda) |kept>, exact, before the cut. */
int sfci_hv(sfci_ctx* ctx, const sfci_vec* in, sfci_vec* out, int64_t nd, double lambda, double coeff,
            double* diag);

/* calc_energy (src/functions.cpp:1272-1313): <v|H|v>. */
int sfci_energy(sfci_ctx* ctx, const sfci_vec* v, double* e);

/* DiagCal_bit (src/bitstr.cpp:433-476): <D|H|D> of one determinant. */
int sfci_diag_element(sfci_ctx* ctx, uint64_t det, double* e);

/* dynVec::dot (src/all.h:317-351). */
int sfci_dot(sfci_ctx* ctx, const sfci_vec* a, const sfci_vec* b, double* out);
/* dynVec::norm (src/all.h:249-258). */
int sfci_norm(sfci_ctx* ctx, const sfci_vec* a, double* out);
/* dynVec::scalarMtply (src/all.h:260-265). */
int sfci_scale(sfci_ctx* ctx, sfci_vec* a, double s);
/* dynVec::addTwo (src/all.h:407-480): v <- top-{v.cap}(v + a*w), zeros kept, key-sorted. */
int sfci_add_two(sfci_ctx* ctx, sfci_vec* v, const sfci_vec* w, double a);
/* dynVec::shrinkFrom (src/all.h:281-299): dst <- top-{dst.cap}(src); src is left as is. */
int sfci_shrink_from(sfci_ctx* ctx, sfci_vec* dst, const sfci_vec* src);
/* No counterpart in the reference.  sfci_shrink_from for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops): the same result, bit for bit,
 * with no hipMalloc/hipFree and no allocating library call.  Its scratch is one fixed
 * 18 KiB device array per GPU (a histogram and per-chunk counts), set up when a context is
 * created; every call on one GPU shares it, so calls that overlap on one GPU race.
 * src longer than SFCI_NOALLOC_MAX_N entries is refused with SFCI_E_CAPACITY. */
#define SFCI_NOALLOC_MAX_N 4294967295LL
int sfci_shrink_from_noalloc(sfci_ctx* ctx, sfci_vec* dst, const sfci_vec* src);
/* Restart recombination (src/sparse_solver.cpp:492-516): out <- top-{out.cap}(sum_i c_i v_i),
 * duplicates merged.  out may be one of the inputs. */
int sfci_lincomb(sfci_ctx* ctx, const sfci_vec* const* vecs, const double* c, int n, sfci_vec* out);

/* ---- Ground-state driver ----------------------------------------------------------
 * sparse_krylov_solve_eigenval (src/sparse_solver.cpp:246-548) in host C++ on top of
 * the calls above: truncated Lanczos with full projection onto the stored Krylov
 * vectors, Nl x Nl generalized eigenproblem, restart from the Ritz vector nearest the
 * previous Ritz value (the first restart takes the lowest), stop when the Ritz value
 * drops by less than delta_e. */
typedef struct {
    int64_t nd;          /* Nd: determinants kept per Krylov vector */
    int nl;              /* Nl: Krylov vectors per restart */
    int num_restart;     /* maximum restarts */
    int64_t hpsi_cap;    /* capacity of H.v results (HPsiCap, src/main.cpp:194): 60,000,000 */
    double delta_e;      /* convergence threshold on the Ritz value (src/sparse_solver.cpp:269): 1e-5 */
    int verbose;         /* 1: print progress lines like the reference's driver */
    double e_prev_init;  /* "previous" Ritz value before the first restart.  The reference uses 0.0
                            (src/sparse_solver.cpp:270), which ends the solve after the first restart
                            whenever the energy is positive; +INFINITY always compares two restarts. */
    double* ms_restart;  /* optional (capacity num_restart): wall time from the start of the solve
                            to the end of each restart (no reference counterpart) */
} sfci_lanczos_params_t;

typedef struct {
    int n_restarts;      /* restarts run */
    double e_final;      /* <v|H|v> of the final vector ("final e:", src/sparse_solver.cpp:545) */
    double ms_total;     /* wall time of the solve */
    double ms_hv;        /* wall time spent in sfci_hv */
    int64_t n_hv;        /* number of sfci_hv calls */
    int n_screened;      /* restarts whose products were screened (sfci_set_screen) */
    int n_reverts;       /* screened restarts rejected by the unscreened energy readout (see sfci_ground_state) */
} sfci_lanczos_result_t;

/* Runs from the single determinant init_det (coefficient 1).  ritz and e_trunc receive
 * one value per restart (capacity num_restart): the selected Ritz value and <v|H|v>
 * after truncation.  v_out (optional, capacity >= nd) receives the final vector.
 * With screening on (sfci_set_screen), the unscreened energy readout <v|H|v> judges every
 * restart and the lowest-energy vector is kept: a screened restart whose energy rises by
 * more than 1e-9 Ha over the best so far is rejected, the best vector is restored and the
 * rest of the solve runs unscreened; a screened solve whose energy drops by less than
 * delta_e hands over to one unscreened restart, which ends the solve unless it lowers the
 * energy by delta_e or more.  The first unscreened restart takes the lowest Ritz value.
 * The solve returns the lowest-energy vector (e_final is its energy).  The context's
 * screening setting is the same on return. */
int sfci_ground_state(sfci_ctx* ctx, uint64_t init_det, const sfci_lanczos_params_t* params, double* ritz,
                      double* e_trunc, sfci_lanczos_result_t* result, sfci_vec* v_out);

/* ---- linear-solver products (the other callers of the H.v boundary) ------------------ */

/* dst = top_{dst.cap}({|c| > cut} of src), key order: the cut-then-cap tail of
 * spmspvm_bucket_linear (src/functions.cpp:1016-1062). */
int sfci_shrink_cut(sfci_ctx* ctx, sfci_vec* dst, const sfci_vec* src, double cut);

/* spmspvm_bucket_linear (src/functions.cpp:858-1094).  w = (H - lambda) in, every distinct
 * target kept (no cut, no cap); hij[j] = <basis[j]|w> for j < n_basis (exact);
 * w -= hij[n-1] basis[n-1]; w -= hij[n-2] basis[n-2] (when n_basis > 1, addTwo keeping
 * zeros); out = top_{out.cap}({|w| > cut}), key order.  basis[n_basis-1] is normally `in`.
 * The reference leaves Hij untouched for an empty input; here hij is zeroed. */
int sfci_hv_linear(sfci_ctx* ctx, const sfci_vec* in, sfci_vec* out, double lambda, const sfci_vec* const* basis,
                   int n_basis, double* hij);

/* spmspvm_bucket_res_norm (src/functions.cpp:1096-1259).  w = (H - lambda) x - b, uncut;
 * *res_norm = ||w|| (exact sum of squares); out = top_{out.cap}({|w| > cut}), key order. */
int sfci_hv_residual(sfci_ctx* ctx, const sfci_vec* x, sfci_vec* out, double lambda, const sfci_vec* b,
                     double* res_norm);

/* dynVec::addThree (src/all.h:482-608): out = top_{out.cap}(a x + b y + c z), then
 * entries with |c| <= 1e-14 dropped, key order.  out may alias an input. */
int sfci_add_three(sfci_ctx* ctx, sfci_vec* out, const sfci_vec* x, const sfci_vec* y, const sfci_vec* z, double a,
                   double b, double c);

typedef struct {
    int64_t nd;          /* capacity of every Krylov vector, of w, r0 and x_m (Nd) */
    int num_iter;        /* Krylov dimension m: m + 1 products per restart (numIteration) */
    int num_restart;     /* restarts (numRestart) */
    int verbose;         /* 1: print the reference's per-restart lines */
} sfci_linear_params_t;

/* sparse_krylov_solve_linear (src/sparse_solver.cpp:561-731): solves (H - target_e) x = b
 * from x0 = 0 with restarted Krylov cycles: K_0 = r0/|r0|; for i = 0..m,
 * w = sfci_hv_linear(K_i, K_0..K_i), Hess(j,i) = Hess(i,j) = hij[j], S(i,j) = <K_j|K_i>,
 * K_{i+1} = w/|w|; y = lstsq(S Hess[:, :m], |r0| S[:, 0]); x_m = top_Nd(x_m + sum_j y_j K_j);
 * r0 = -top_Nd(cut((H - target_e) x_m - b)).  residual[r] receives the residual norm after
 * restart r (capacity num_restart); x_out (capacity >= nd) the solution. */
int sfci_linear_solve(sfci_ctx* ctx, const sfci_vec* b, double target_e, const sfci_linear_params_t* params,
                      sfci_vec* x_out, double* residual);

/* sparse_solve_SI (src/sparse_solver.cpp:733-772): one shift-invert step,
 * v_out = x/|x| for x = sfci_linear_solve(b, target_e), *energy = <v_out|H|v_out>. */
int sfci_shift_invert(sfci_ctx* ctx, const sfci_vec* b, double target_e, const sfci_linear_params_t* params,
                      sfci_vec* v_out, double* energy, double* residual);

/* minres_sp (src/minres.cpp:73-225): MINRES on (H - lambda) x = rhs from `guess`, every
 * vector capped at nd by the reference's products (sfci_hv with coeff -1 / +1, addTwo,
 * addThree).  Stops after `iters` steps or when the running residual estimate drops below
 * `tol` (compared with the squared estimate, as the reference does).  *iters_done and
 * *res2 (optional) receive the step count and the final squared estimate. */
int sfci_minres(sfci_ctx* ctx, const sfci_vec* rhs, const sfci_vec* guess, sfci_vec* result, int64_t nd, int iters,
                double tol, double lambda, int* iters_done, double* res2);

/* calc_CISD (src/sparse_solver.cpp:15-114): the CISD block of ref_det (ref_det and every
 * determinant coupled to it by a nonzero element), its dense Hamiltonian, eigenpair
 * `state_idx` (0 = lowest).  out (capacity >= block size) receives the eigenvector in
 * key order, normalized, its largest-|c| component (smallest key among ties) positive;
 * *eigenvalue the eigenvalue; *block_size (optional) the block dimension. */
int sfci_cisd(sfci_ctx* ctx, uint64_t ref_det, int state_idx, sfci_vec* out, double* eigenvalue,
              int64_t* block_size);

/* sparse_main (src/sparse_solver.cpp:774-809): CISD state 2 of init_det, normalized, then
 * sfci_shift_invert at target_e.  (The reference's task=excited branch leaves this call
 * commented out, src/main.cpp:235-242.) */
int sfci_excited(sfci_ctx* ctx, uint64_t init_det, double target_e, const sfci_linear_params_t* params,
                 sfci_vec* v_out, double* energy, double* e_cisd);

/* ---- learned candidate screener ------------------------------------------------------
 * No reference counterpart: the reference keeps every candidate of H.v
 * (src/functions.cpp:623-639).  With screening on, sfci_hv (and every driver built on it)
 *   1. emits only connections with |H_ij coeff c_j| > eps_pre (diagonal tags always),
 *   2. scores each emitted candidate (mode 2: the network's predicted log|c| of its target;
 *      mode 1: |H_ij coeff c_j| alone),
 *   3. keeps the candidates whose score reaches the keep-th largest score
 *      (keep = keep_factor * nd; ties at that score kept) and every diagonal tag,
 *   4. merges the kept candidates exactly and applies the cut, cap and order of sfci_hv.
 * Values are FP32 products as in SFCI_PREC_MIXED; sums stay exact.  The diag output is
 * the exact <kept|y> of the screened y. */
typedef struct sfci_mlp sfci_mlp;
#define SFCI_MLP_INPUTS 66     /* +-1 occupation of 64 spin orbitals, excitation rank, plog2|v| */
#define SFCI_MLP_HIDDEN 256
/* flat FP32 parameters: W1[256][66] (row = hidden unit), b1[256], W2[256][256],
 * b2[256], w3[256], b3 */
#define SFCI_MLP_NPARAM 83201

/* One training sample: a candidate's target determinant, its excitation rank, plog2 of its
 * value |H_ij coeff c_j| (the FP32 exponent plus the mantissa fraction: log2 piecewise
 * linear, -150 for zero), and the label log|y_t| (natural log) of its target in the full
 * product. */
typedef struct {
    uint64_t det;
    float logv;
    float label;
    uint32_t cls;
    uint32_t pad;
} sfci_screen_sample_t;

/* Network on the context's GPU; weights uniform in +-1/sqrt(fan_in) from `seed`, zero biases. */
int sfci_mlp_create(sfci_ctx* ctx, uint64_t seed, sfci_mlp** out);
void sfci_mlp_destroy(sfci_mlp* net);
/* Master FP32 parameters (SFCI_MLP_NPARAM floats); set resets the Adam state. */
int sfci_mlp_set_params(sfci_mlp* net, const float* params);
int sfci_mlp_get_params(const sfci_mlp* net, float* params);
/* Input and label normalisation: feature 65 = (logv - x_shift) * x_scale (FP32); training target
 * = (label - y_shift) * y_scale (scales > 0). */
int sfci_mlp_set_norm(sfci_mlp* net, float x_shift, float x_scale, float y_shift, float y_scale);
/* y[i] = network(x[i]) for explicit features x[n][66] (rounded to BF16 on the GPU). */
int sfci_mlp_forward(sfci_mlp* net, const float* x, int64_t n, float* y);
/* Gradient of the mean squared error mean_i (y_i - t_i)^2 over x[n][66] (n % 256 == 0)
 * in the flat parameter layout; *loss (optional) receives the error. */
int sfci_mlp_gradient(sfci_mlp* net, const float* x, const float* t, int64_t n, float* grad, float* loss);
/* One Adam step (bias-corrected) from a given flat gradient. */
int sfci_mlp_adam_step(sfci_mlp* net, const float* grad, float lr, float beta1, float beta2, float eps);
/* `steps` Adam steps (lr, 0.9, 0.999, 1e-8) on batches of `batch` samples (batch % 256 == 0)
 * drawn from samples[n] by a hash of (seed, step, row); loss_curve (optional, `steps` floats)
 * receives each batch's mean squared error in normalised units. */
int sfci_mlp_train(sfci_mlp* net, const sfci_screen_sample_t* samples, int64_t n, int steps, int batch, float lr,
                   uint64_t seed, float* loss_curve);
/* Network output (normalised units) for samples, and the features it sees (n x 66). */
int sfci_mlp_predict_samples(sfci_mlp* net, const sfci_screen_sample_t* samples, int64_t n, float* y);
int sfci_mlp_sample_features(sfci_mlp* net, const sfci_screen_sample_t* samples, int64_t n, float* x);

#define SFCI_SCREEN_OFF 0
#define SFCI_SCREEN_MAGNITUDE 1   /* rank candidates by |H_ij coeff c_j| */
#define SFCI_SCREEN_NETWORK 2     /* rank candidates by the network's prediction */
#define SFCI_SCREEN_NETWORK_MAGNITUDE 3   /* rank by the network's predicted ln|y| of the target plus ln|H_ij c_j| */
/* Screening of sfci_hv (see above); net is required for SFCI_SCREEN_NETWORK and must
 * outlive its use. */
int sfci_set_screen(sfci_ctx* ctx, int mode, sfci_mlp* net, double keep_factor, double eps_pre);
/* The current screening setting (any output pointer may be null). */
int sfci_get_screen(const sfci_ctx* ctx, int* mode, sfci_mlp** net, double* keep_factor, double* eps_pre);
/* Training samples from the product H v: the candidates of every entry of v that pass eps_pre
 * (off-diagonal only), a deterministic subsample of at most max_samples of them chosen by
 * a hash of (seed, candidate), each labelled with log|y_t| of its target in the full,
 * uncut y = H v (log of 1e-30 for an exact zero).  *n_out receives the count. */
int sfci_screen_samples(sfci_ctx* ctx, const sfci_vec* v, double eps_pre, int64_t max_samples, uint64_t seed,
                        sfci_screen_sample_t* out, int64_t* n_out);

/* ---- 128-bit determinant keys: more than 64 spin orbitals (up to 128) -------------
 * No reference counterpart: the reference stores a determinant in one 64-bit word
 * (dynVec<u64,double>, src/all.h:150-194) and stops at 64 spin orbitals
 * (src/main.cpp:84-85).  These entry points extend SpMVMtply_bucket
 * (src/functions.cpp:511-804) to keys of two words.  A key is stored as two uint64:
 * word 0 = spin orbitals 0-63, word 1 = spin orbitals 64-127; vectors are sorted
 * ascending by the 128-bit value word1:word0.  The context takes the same _IN files as
 * sfci_create; every other convention (return codes, exact sums, tie rule, cut) is that
 * of the 64-bit entry points. */
typedef struct sfci_wctx sfci_wctx;
typedef struct sfci_wvec sfci_wvec;

/* n_so even <= 128, n_elec even < 64, n_elec / 2 <= n_so / 2. */
int sfci_wide_create(const char* in_path, int n_so, int n_elec, int device, sfci_wctx** out);
void sfci_wide_destroy(sfci_wctx* ctx);
const char* sfci_wide_last_error(const sfci_wctx* ctx);
int sfci_wide_info(const sfci_wctx* ctx, sfci_info_t* info);
int sfci_wide_set_cut(sfci_wctx* ctx, double cut);
int sfci_wide_set_precision(sfci_wctx* ctx, int precision);
/* Device bytes the candidate buffers of one pass may use (about 52 bytes per candidate
 * plus a fixed 420 MB of partition slack). */
int sfci_wide_set_workspace(sfci_wctx* ctx, uint64_t bytes);
int sfci_wide_hv_stats(const sfci_wctx* ctx, sfci_hv_stats_t* st);
/* Tuning and test knobs of the 128-bit merge; results are identical for every value.
 * hash_bits: bits of the key hash that bucket candidate records (9..16; 0 = from the pass
 * size, the default).  table_fill: occupied slots (1..2867) at which a bucket's LDS table
 * counts as full and the bucket is merged in hash classes (0 = 2867, the default; small
 * values force class splits). */
int sfci_wide_set_merge(sfci_wctx* ctx, int hash_bits, int table_fill);
int sfci_wide_vec_create(sfci_wctx* ctx, int64_t cap, sfci_wvec** out);
void sfci_wide_vec_destroy(sfci_wvec* v);
int64_t sfci_wide_vec_len(const sfci_wvec* v);
/* keys: 2 * n words (entry i = keys[2i], keys[2i+1]), strictly ascending. */
int sfci_wide_vec_upload(sfci_wvec* v, const uint64_t* keys, const double* vals, int64_t n);
int sfci_wide_vec_download(const sfci_wvec* v, uint64_t* keys, double* vals, int64_t max_n, int64_t* n_out);
/* sfci_hv on 128-bit keys. */
int sfci_wide_hv(sfci_wctx* ctx, const sfci_wvec* in, sfci_wvec* out, int64_t nd, double lambda, double coeff,
                 double* diag);
/* The vector operations of the ground-state driver on 128-bit keys, with the semantics of
 * sfci_vec_copy, sfci_vec_clear, sfci_dot, sfci_norm, sfci_scale, sfci_add_two,
 * sfci_shrink_from and sfci_lincomb (dynVec, src/all.h:249-480,610-623; restart
 * recombination, src/sparse_solver.cpp:492-516). */
int sfci_wide_vec_copy(sfci_wvec* dst, const sfci_wvec* src);
int sfci_wide_vec_clear(sfci_wvec* v);
int sfci_wide_dot(sfci_wctx* ctx, const sfci_wvec* a, const sfci_wvec* b, double* out);
int sfci_wide_norm(sfci_wctx* ctx, const sfci_wvec* a, double* out);
int sfci_wide_scale(sfci_wctx* ctx, sfci_wvec* a, double s);
int sfci_wide_add_two(sfci_wctx* ctx, sfci_wvec* v, const sfci_wvec* w, double a);
int sfci_wide_shrink_from(sfci_wctx* ctx, sfci_wvec* dst, const sfci_wvec* src);
int sfci_wide_lincomb(sfci_wctx* ctx, const sfci_wvec* const* vecs, const double* c, int n, sfci_wvec* out);
/* <v|H|v> (calc_energy, src/functions.cpp:1272-1313) as the exact diag of H.v over all of v:
 * sums exact, rounded per target value and once for the total (the 64-bit sfci_energy
 * rounds per source instead; the two agree to the last few bits). */
int sfci_wide_energy(sfci_wctx* ctx, const sfci_wvec* v, double* e);
int sfci_wide_sync(sfci_wctx* ctx);
/* sparse_krylov_solve_eigenval (src/sparse_solver.cpp:246-548) on 128-bit keys: the
 * driver of sfci_ground_state without candidate screening.  init_det is two words
 * (init_det[0] = spin orbitals 0-63). */
int sfci_wide_ground_state(sfci_wctx* ctx, const uint64_t* init_det, const sfci_lanczos_params_t* params,
                           double* ritz, double* e_trunc, sfci_lanczos_result_t* result, sfci_wvec* v_out);

#ifdef __cplusplus
}
#endif

#endif /* SFCI_GPU_H */
