/*
 * oops_expv.h — C ABI of the MI355X-native CCE Krylov-expv hot path.
 *
 * Plain C types only (no torch, no HIP types).  Every call returns an int
 * status (OOPS_OK == 0, negative on error; oops_last_error() gives the text)
 * instead of stopping the process the way the reference's Fortran does.
 * Complex numbers are passed as interleaved (re, im) doubles, i.e. the
 * memory layout of std::complex<double> / double _Complex.
 *
 * Reference interfaces replaced (paths relative to the reference's src/):
 *   oops_term_table        <- the flattened term arrays MatExpVector builds
 *                             (source/math/MatExp.cpp:219-262) and hands to
 *                             main_mkl_/main_cache_ (expv/include/main_cache.h:6)
 *   oops_hamvec            <- hamvec_cuda3_ / hamvec_func3_
 *                             (expv/source/hamvec_cuda3.cu:661-665,
 *                              expv/source/hamvec_func3.cpp:399)
 *   oops_expv / main_cache_ <- MAIN_CACHE / MAIN_MKL (expv/source/main_cache.f:1-136,
 *                             expv/source/main_mkl.f:1-129) as called by
 *                             MatExpVector::runInexplicitGPU / runInexplicit
 *                             (source/math/MatExp.cpp:216-521)
 *   oops_hahn_batch_*      <- no reference counterpart: the reference calls one
 *                             expv per cluster from CCE::run_each_clusters
 *                             (source/app/cce.cpp:54-70); the batch runs every
 *                             cluster's two Hahn-echo sweeps on the device and
 *                             returns L_c(t) instead of the n x nt state matrix.
 *   oops_cce_*             <- the CCE host pipeline: NV eigenstates
 *                             (source/app/DefectCenter.cpp:32-45), per-cluster
 *                             Liouvillians (source/app/ensemble_cce.cpp:33-113),
 *                             cluster enumeration (source/spin/SpinCluster/
 *                             SpinClusterAlgorithm.cpp:98-148) and the CCE
 *                             reduction (source/app/cce.cpp:151-189).
 */
#ifndef OOPS_EXPV_H
#define OOPS_EXPV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OOPS_OK 0
#define OOPS_ERR_ARG (-1)
#define OOPS_ERR_HIP (-2)
#define OOPS_ERR_NOMEM (-3)
#define OOPS_ERR_UNSUPPORTED (-4)
#define OOPS_ERR_NUMERIC (-5)
#define OOPS_ERR_CAPACITY (-6)

typedef struct oops_ctx oops_ctx;
typedef struct oops_hahn_batch oops_hahn_batch;

/* One operator H = sum_t coeff[t] * kron_b mats[t,b] in the reference's
 * flattened layout (MatExp.cpp:228-262): term t has nbody[t] bodies at
 * pos[pos_offset[t] .. pos_offset[t+1]), body b's local matrix is the
 * dims[b] x dims[b] column-major block mats[mat_offset[b] ..], spin 0 is the
 * slowest Kronecker factor, spin_dim[s] is the local dimension of spin s. */
typedef struct {
    int32_t nspin;
    int32_t nterm;
    const double* coeff;        /* [nterm] real multipliers */
    const int64_t* nbody;       /* [nterm] 0, 1 or 2 */
    const int64_t* pos_offset;  /* [nterm + 1] */
    const int64_t* pos;         /* [pos_offset[nterm]] */
    const int64_t* dims;        /* [pos_offset[nterm]] */
    const int64_t* mat_offset;  /* [pos_offset[nterm] + 1] (complex entries) */
    const double* mats;         /* interleaved complex, [mat_offset[last]] */
    const int64_t* spin_dim;    /* [nspin] */
} oops_term_table;

/* Flip-mask (slot) form of an operator, computed on the host: with
 * power-of-two local dimensions every entry of a merged local operator moves
 * element i to i ^ mask, so (H x)_i = sum_s coef[s][i] x[i ^ masks[s]].
 * Slots [0, nslot64) change at most one spin's digits (slot 0 is mask 0),
 * slots [nslot64, nslot64 + nslot32) change two.  With masks or coef NULL
 * only the counts are returned; otherwise masks gets nslot64 + nslot32 values
 * and coef (nslot64 + nslot32) x n complex values, slot-major.  No reference
 * counterpart: it is the layout the mixed Lanczos kernel builds on the device
 * (OOPS_ERR_CAPACITY beyond max_slots). */
int oops_slot_form(const oops_term_table* T, int32_t max_slots, int32_t* nslot64, int32_t* nslot32, int32_t* masks,
                   double* coef);

/* Pass plan of the tiled product the large path uses for Kronecker
 * operators of 2^19 elements and more (*npasses = 0: the flat ELL product).
 * mode 1 with v, w: w = H v on the host through that plan (the kernels'
 * tile-to-index mapping, checked to be a bijection in every pass: status
 * OOPS_ERR_NUMERIC otherwise); mode 2: w = H v through the flat ELL tables;
 * mode 3: w = H v through the two-axis product the large path uses for
 * spin-1/2 Liouvillians of 8-12 spins whose terms act on ket digits only or
 * on bra digits only (*npasses = 2; OOPS_ERR_UNSUPPORTED when the table does
 * not split), with the kernels' tile mapping checked to be a bijection.
 * No reference counterpart (host check of the device layout). */
int oops_tile_plan(const oops_term_table* T, int32_t* npasses, int32_t mode, const double* v, double* w);

/* --- context ------------------------------------------------------------ */
int oops_ctx_create(int32_t device, oops_ctx** out);
int oops_ctx_destroy(oops_ctx* ctx);
const char* oops_last_error(const oops_ctx* ctx);
/* Device properties: number of compute units, LDS bytes per workgroup. */
int oops_ctx_info(const oops_ctx* ctx, int32_t* num_cu, int32_t* lds_bytes, char* name, int32_t name_len);
/* Launch shape of the persistent batch kernels: wg_per_cu workgroups per
 * compute unit for kernel kind `kind` (the kinds oops_hahn_batch_timing
 * reports: 0-4, 6-9) in batches created on ctx after the call; 0 restores
 * the default (the number one CU holds: 4 for kind 0, 2 for 1-3, 6 and 9,
 * 1 for 4, 7 and 8).  Work units pull from a device queue, so results do
 * not depend on it.  No reference counterpart (the autotuner's knob). */
int oops_ctx_set_launch(oops_ctx* ctx, int32_t kind, int32_t wg_per_cu);

/* --- single operator ---------------------------------------------------- */
/* w = H v for host vectors of length prod(spin_dim) (hamvec_cuda3_).
 * Local dimensions need not be powers of two (spin-1: 3, its Liouville
 * space: 9, mixed kinds; hamvec_func3.cpp:307-389 takes any): such an
 * operator is embedded in the next power of two of its largest local
 * dimension on the device and the vectors are mapped to and from the
 * caller's index order on the host.  The same holds for oops_expv and
 * oops_expv_ex; oops_hamvec_noalloc, whose vectors stay on the device,
 * returns OOPS_ERR_UNSUPPORTED for them. */
int oops_hamvec(oops_ctx* ctx, const oops_term_table* H, const double* v, double* w);
/* y = H x like oops_hamvec, but x and y are DEVICE pointers (prod(spin_dim)
 * complex doubles each, owned by the caller) and the call allocates no device
 * memory: the compiled operator tables are copied from the host into the
 * library's fixed 16 MiB device array on every call (one array per device,
 * set up by oops_ctx_create) and the product runs from there (flat form, or
 * the tiled form for states of 2^19 elements and more).  Power-of-two local
 * dimensions only (OOPS_ERR_UNSUPPORTED otherwise).  Tables beyond
 * 16 MiB return OOPS_ERR_CAPACITY.  The call returns after the product has
 * finished.  Calls that overlap on one device (several threads, several
 * contexts) share the array and race.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). */
int oops_hamvec_noalloc(oops_ctx* ctx, const oops_term_table* H, const double* x, double* y);

/* Multi-output Krylov expv of one state: w(t_k) = exp(-i H t_k) v for the
 * ascending times ta[0..nt) (MAIN_CACHE / zgexpv_cache: Lanczos(klim) norm
 * estimate from DLARAN seed (0,0,0,5), Krylov dimension m (2..40, < n),
 * tolerance tol per unit time).  The basis (n x (m+1) complex) stays in
 * device memory; only scalars cross to the host per Krylov step.
 * out_mode 0: out = w_seq, n x nt complex column-major (the reference's
 * w_seq); out_mode 1: out = <v, w(t_k)>, nt complex.
 * stats (may be NULL): [0] products, [1] small exponentials, [2] squarings,
 * [3] steps, [4] rejections, [5] breakdown flag, [6] norm estimate,
 * [7] device time of the Krylov steps in ms. */
int oops_expv(oops_ctx* ctx, const oops_term_table* H, const double* v, int32_t nt, const double* ta, int32_t klim,
              int32_t m, double tol, int32_t out_mode, double* out, double* stats);

/* oops_expv with options in `precision`.  Bit 0: the basis stored in complex
 * FP32 (half the HBM traffic of the Gram-Schmidt passes; products, sums and
 * the small exponential stay FP64); clear: FP64 throughout (0 = oops_expv).
 * Bit 1: output times inside steps.  The steps run to min(t_new, ta[nt-1])
 * under Expokit's error control instead of ending at every output time
 * (zgexpv_cache.f:418-440), and every output time inside a step is evaluated
 * from that step's basis.  Fewer Krylov cycles on grids finer than the
 * natural step; results agree with the reference's schedule to the
 * tolerance, not step for step.  Other bits: OOPS_ERR_ARG. */
int oops_expv_ex(oops_ctx* ctx, const oops_term_table* H, const double* v, int32_t nt, const double* ta,
                 int32_t klim, int32_t m, double tol, int32_t precision, int32_t out_mode, double* out,
                 double* stats);

/* Fortran-ABI drop-in for the reference's main_cache_ (main_cache.h:6),
 * same argument list; runs oops_expv with out_mode 0 on device 0. */
void main_cache_(size_t* nspin, size_t* nTerm, double* coeff_lst, size_t* nbody_lst, size_t* pos_i_idx,
                 size_t* pos_i_lst, size_t* dim_i_lst, size_t* mat_i_idx, double* mat_i_lst, size_t* ham_dim,
                 size_t* nspin_dim, double* v, size_t* pos_idx, size_t* mat_idx, size_t* k,
                 size_t* maxThreadsPerBlock, size_t* maxGridSize, size_t* tn, double* ta, size_t* m, double* tol,
                 size_t* itrace, double* w_seq, size_t* w_seq_len);

/* --- link-compatible symbols of the reference's expv library ------------------
 * Same names and argument order as expv/include/{main_mkl,krylov_expv,
 * expokit}.h and hamvec_func3.cpp:399 / hamvec_cuda3.cu:661-665; complex
 * arrays are interleaved doubles; they run on device 0. */
void main_mkl_(size_t* nspin, size_t* nTerm, double* coeff_lst, size_t* nbody_lst, size_t* pos_i_idx,
               size_t* pos_i_lst, size_t* dim_i_lst, size_t* mat_i_idx, double* mat_i_lst, size_t* ham_dim,
               size_t* nspin_dim, double* v, size_t* pos_idx, size_t* mat_idx, size_t* k, size_t* tn, double* ta,
               size_t* m, double* tol, size_t* itrace, double* w_seq, size_t* w_seq_len);
/* w_seq(:, k) = exp(ta[k] A) v, A dense n x n column-major (krylov_zgexpv) or
 * 1-based COO (krylov_zcooexpv).  Returns 0 or a negative OOPS_ERR_*. */
int krylov_zgexpv(const int n, const double* a, const double* v, const int tn, const double* ta, double* w_seq,
                  const int klim, const int m, const double tol, const int itrace);
int krylov_zcooexpv(const int n, const int nz, const int* ia, const int* ja, const double* a, const double* v,
                    const int tn, const double* ta, double* w_seq, const int klim, const int m, const double tol,
                    const int itrace);
/* w = sum_t coeff_zplx[t] (x)_b A_tb v (host vectors; strides and w_med unused). */
void hamvec_func3_(int* nspin, int* nTerm, double* coeff_lst_zplx, size_t* nbody_lst, size_t* pos_i_idx,
                   size_t* pos_i_lst, size_t* dim_i_lst, size_t* mat_i_idx, double* mat_i_lst, size_t* ham_dim,
                   size_t* nspin_dim, size_t* nspin_m_lst, size_t* nspin_n_lst, double* v, double* w, double* w_med);
/* Handle = an oops_ctx on device 0; device pointers passed as size_t. */
void hamvec_cuda3_init_(size_t* handle);
void hamvec_cuda3_term_(size_t* handle);
void hamvec_cuda3_(size_t* handle, int* nspin, int* nTerm, double* coeff_lst_zplx, size_t* nbody_lst,
                   size_t* pos_i_idx, size_t* pos_i_lst, size_t* dim_i_lst, size_t* mat_i_idx,
                   size_t* dev_mat_i_lst, size_t* ham_dim, size_t* nspin_dim, size_t* nspin_m_lst,
                   size_t* nspin_n_lst, size_t* dev_v, size_t* dev_w, size_t* dev_w_med, size_t* dev_coeff,
                   size_t* maxThreadsPerBlock, size_t* maxGridSize);
/* exp(t H) (m x m, ideg = 6) into wsp(1..m*m), *iexph = 1, *ns squarings as
 * ZGPADM counts them.  m <= 42 runs in one workgroup; larger m (the dense
 * propagators of MatExp::pade_exp_mat) on rocBLAS products and a rocSOLVER
 * LU solve, loaded at first use (*iflag = -2 when they are missing). */
void zgpadm_(int* ideg, int* m, double* t, double* H, int* ldh, double* wsp, int* lwsp, int* ipiv, int* iexph,
             int* ns, int* iflag);

/* --- batched Hahn echo ---------------------------------------------------- */
/* One cluster: opA generates the ket sweep exp(-i opA tau) u0, opB the bra
 * sweep exp(-i opB tau) y0 (pass -lvB^dagger for lvB), and
 *   L(t_k) = Re( scale * < bra(tau_k), ket(tau_k) > ). */
typedef struct {
    const oops_term_table* opA;
    const oops_term_table* opB;
    const double* u0;  /* n complex */
    const double* y0;  /* n complex */
    double scale;
} oops_hahn_item;

/* Compile every item's operators, bucket the items by state dimension and
 * upload everything to device memory.  m: Krylov dimension (20 or 30);
 * clusters with n <= m use the dense exponential.  klim: Lanczos steps of the
 * norm estimate (the reference default is 10), tol: Expokit tolerance.
 * precision 0: complex FP64 throughout (the reference's Arnoldi + Pade).
 * 1: mixed — 2-body couplings applied in complex FP32 and the basis stored in
 * FP32; 1-body terms, the recurrence coefficients, re-orthogonalisation and
 * the small exponential in FP64.  Buckets whose generators are all Hermitian
 * run Lanczos with a Chebyshev exponential of the tridiagonal matrix (in the
 * flip-mask form when every operator has at most 9 one-spin and 12 two-spin
 * flip masks: all products of 4-spin spin-1/2 Liouvillians), the others
 * Arnoldi + Pade.  2: mixed, always Arnoldi + Pade.  3: as 1 without the
 * flip-mask form (ELL product in the Lanczos kernel).  4: FP64 throughout
 * with the mixed path's algorithm for the buckets the flip-mask Lanczos
 * kernel takes (Lanczos, Chebyshev exponential, output times inside steps,
 * the same tolerance), the FP64 paths of precision 0 for the others: it
 * isolates what the FP32 parts of precision 1 gain. */
int oops_hahn_batch_create(oops_ctx* ctx, int32_t nitems, const oops_hahn_item* items, int32_t m, int32_t klim,
                           double tol, int32_t precision, oops_hahn_batch** out);
/* Run every bucket for the ascending sweep times tau[0..nt) (device-resident
 * results).  Blocks until the device is done. */
int oops_hahn_batch_run(oops_hahn_batch* b, int32_t nt, const double* tau);
/* Copy results: L [nitems x nt] (row per item, item order of create), stats
 * [nitems x 24] int32 per-cluster counters: [0..3] steps, products, rejections,
 * small exponentials of sweep A, [4..7] the same for sweep B, [8] failure flags,
 * [9] exponential substeps, [10..13] the two norm estimates (doubles).  Either may be NULL.
 * L of a failed cluster (flags & 0xFF nonzero) is NaN where its sweeps
 * stopped before writing it (result buffers start as NaN). */
int oops_hahn_batch_results(oops_hahn_batch* b, double* L, int32_t* stats);
/* --- pulse sequences with per-item time grids ------------------------------
 * One cluster under a Hahn echo (npulse = 1) or CPMG-N (npulse = N >= 2,
 * Pulse_Interval / riffle of misc.cpp:59-84: segment fractions 1/(2N), 1/N,
 * ..., 1/N, 1/(2N) with generators lvA, lvB, lvA, ...).  Hahn echo: t holds
 * the sweep times t/2, L(t) = Re(scale <exp(-i opB t/2) y0, exp(-i opA t/2) u0>)
 * as for oops_hahn_item.  CPMG-N: t holds the times t and
 * L(t) = Re(scale <y0, U_N .. U_0 u0>), U_j = exp(-i L_j f_j t); opC = lvB and
 * opD = -lvA^dagger are required.  CPMG-N runs on the dense path (n <= m) or
 * on the mixed Lanczos path (precision 1, Hermitian generators). */
typedef struct {
    const oops_term_table* opA;  /* lvA */
    const oops_term_table* opB;  /* -lvB^dagger */
    const oops_term_table* opC;  /* lvB (CPMG-N) or NULL */
    const oops_term_table* opD;  /* -lvA^dagger (CPMG-N) or NULL */
    const double* u0;
    const double* y0;
    double scale;
    int32_t npulse;
    int32_t nt;
    const double* t;             /* nt ascending times >= 0 */
} oops_echo_item;

/* Like oops_hahn_batch_create, with each item's times fixed at creation.  The
 * batch owns a HIP stream, so several batches run concurrently on one device.
 * Results: oops_hahn_batch_results with L concatenated in item order (item i
 * at offset nt_0 + .. + nt_{i-1}). */
int oops_echo_batch_create(oops_ctx* ctx, int32_t nitems, const oops_echo_item* items, int32_t m, int32_t klim,
                           double tol, int32_t precision, oops_hahn_batch** out);
/* Capture the batch's kernel launches into a HIP graph; later launches
 * replay it (one submission for all buckets).  Not for batches with states
 * above 256 elements. */
int oops_echo_batch_capture(oops_hahn_batch* b);
/* Asynchronous launch on the batch's stream; query without blocking; wait. */
int oops_echo_batch_launch(oops_hahn_batch* b);
int oops_echo_batch_query(oops_hahn_batch* b, int32_t* done);
int oops_echo_batch_wait(oops_hahn_batch* b);
/* launch + wait */
int oops_echo_batch_run(oops_hahn_batch* b);
/* Device bytes held by the batch (its admission footprint). */
int oops_echo_batch_footprint(const oops_hahn_batch* b, int64_t* bytes);

/* Per-bucket timing of the last run: for bucket k, its state dimension,
 * item count and kernel time in ms (HIP events on the launch stream). */
int oops_hahn_batch_timing(const oops_hahn_batch* b, int32_t max_buckets, int32_t* n, int32_t* count, double* ms,
                           int32_t* nbuckets);
/* Device time of the last run (ms): one interval from the first launch to
 * the completion of every bucket.  Buckets run concurrently on their own
 * streams, so this is less than the sum of the per-bucket times. */
int oops_hahn_batch_elapsed(const oops_hahn_batch* b, double* ms);
/* Kernel each bucket ran with in the last run: 0 dense VALU Pade (n <= m),
 * 1 Arnoldi FP64, 2 Arnoldi mixed, 3 Lanczos mixed, 4 dense FP64-MFMA
 * propagator (n = 64), 5 single-state path (n > 256), 6 Lanczos mixed in the
 * flip-mask (slot) form, 7 Lanczos mixed for n = 1024 / 4096 (one 1024-thread
 * workgroup per cluster), 8 Lanczos FP64 in the flip-mask form, 9 factored
 * FP64 propagator powers (CPMG-N of unpolarised 4-spin clusters on a uniform
 * grid, built by oops_cce_echo_batch_create); -1 not run. */
int oops_hahn_batch_kernels(const oops_hahn_batch* b, int32_t max_buckets, int32_t* kind);
int oops_hahn_batch_destroy(oops_hahn_batch* b);

/* DLARAN (LAPACK 48-bit multiplicative congruential generator, the Lanczos
 * start vector source of expv/source/expokit/DLARAN.f): draws count values
 * starting from iseed[4] (12-bit limbs, most significant first), updating it. */
int oops_dlaran_fill(int32_t* iseed, int64_t count, double* out);

#ifdef __cplusplus
}
#endif

#endif /* OOPS_EXPV_H */
