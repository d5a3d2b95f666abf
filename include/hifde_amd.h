/*
 * hifde_amd.h -- C ABI of the MI355X (gfx950) HIF-DE factorization library.
 *
 * The reference (arbenson/HIFDE3D) exposes its hot path only as a C++
 * template API (include/Factor.hpp, include/InterpDecomp.hpp,
 * include/Schur.hpp).  This ABI is what that API binds to when it runs on the
 * GPU; include/hifde3d_amd.hpp is the C++ drop-in built on top of it.  Each
 * entry point names the reference interface it replaces.
 *
 * Conventions
 *   - Grid: N interior points per side, NC = N + 1 stored points per side,
 *     DOF (i, j, k) has linear index i + NC*j + NC*NC*k; planes i/j/k = 0 are
 *     the zero Dirichlet boundary (src/Factor.cpp:78-89).
 *   - Dense blocks are column-major.  Indices are 0-based int.
 *   - Status: 0 ok; > 0 a singular pivot was met (LAPACK-style info, the
 *     1-based pivot index); < 0 an error (see HIFDE_ERR_*).  No entry point
 *     throws; hifde_last_error() describes the last failure on this thread.
 *   - Host pointers are read or written only during the call.  Device
 *     pointers (the *_device entry points) must be device memory of the
 *     current HIP device; work is ordered on the factor's own stream and the
 *     call returns after it completes.
 *   - One host thread per factor object at a time.  Distinct objects may be
 *     driven from distinct threads at once: each has its own stream and
 *     buffers, and no entry point synchronizes the whole device or the
 *     legacy default stream (unless that is the stream named by
 *     hifde_set_stream).  A device-wide synchronize issued by the caller
 *     while another thread's call is capturing a graph is refused by HIP
 *     (hipErrorStreamCaptureUnsupported) in the caller's thread; the library
 *     call falls back to eager sweeps.
 */
#ifndef HIFDE_AMD_H_
#define HIFDE_AMD_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hifde_factor hifde_factor;

/* precision of the factorization */
#define HIFDE_FP64 0  /* every block in FP64 (the reference's HIFFactor<double>) */
#define HIFDE_MIXED 1 /* Schur/LU blocks and stored factors in FP32 (every
                         level when epsilon >= 1e-5, else all but the top
                         two; option "fp32_levels"); ID pivots, rank
                         decisions and the level matrix in FP64 */

/* flags */
#define HIFDE_FLAG_FRONT_QUIRK 1 /* reproduce the FRONT-face guard of
                                    src/Factor.cpp:486 (tests k == 0) */

#define HIFDE_OK 0
#define HIFDE_ERR_ARG (-1)
#define HIFDE_ERR_HIP (-2)
#define HIFDE_ERR_SIZE (-3)
#define HIFDE_ERR_STATE (-4)

/* HIFFactor<Scalar>(N, P, epsilon) -- include/Factor.hpp:12-13.
 * Requires NC/P to be a power of two (src/Factor.cpp:36,51-53) and
 * NC = N + 1 < 1024 (HIFDE_ERR_SIZE otherwise; 1023^3 unknowns is far past
 * one GPU's memory for this method). */
int hifde_create(int N, int P, double epsilon, int precision, int flags, hifde_factor** out);
/* ~HIFFactor -- include/Factor.hpp:15 */
void hifde_destroy(hifde_factor* f);
const char* hifde_last_error(void);

/* sp_matrix() filled through Sparse::Add(i, jidx, vals)
 * (include/hifde3d/core/sparse.hpp:103-115): A(rows[t], cols[t]) += vals[t],
 * duplicates summed in the order given. */
int hifde_set_matrix_coo(hifde_factor* f, long long nnz, const int* rows, const int* cols,
                         const double* vals);
/* SetupStencil(factor, N, h, A, V) -- include/setup_stencil.hpp:6-76.
 * a_nodal, v_nodal: (N+2)^3 doubles, index i + (N+2)*(j + (N+2)*k)
 * (include/numtns.hpp:99), h = 1/(N+1).  symmetric == 0 is the reference's
 * stencil (off-diagonal -a(q)/h^2); symmetric == 1 uses -(a(p)+a(q))/(2h^2)
 * with the matching diagonal, identical to the reference for constant a. */
int hifde_set_stencil(hifde_factor* f, const double* a_nodal, const double* v_nodal,
                      int symmetric);

/* SetupStencil as hifde_set_stencil, from nodal tensors in device memory
 * (d_a, d_v: (N+2)^3 doubles each, same layout): the CSR operator is built
 * in HBM with the same rows, entry order and values (bitwise) as the host
 * assembly.  Ordered after the caller's stream (hifde_set_stream); returns
 * when done. */
int hifde_set_stencil_device(hifde_factor* f, const double* d_a, const double* d_v,
                             int symmetric);

/* The assembled operator as CSR (the reference's sp_matrix() before
 * Factor() consumes it; it is kept here): rowptr ((N+1)^3 + 1 ints), colind
 * and vals (nnz each), rows in linear order, entries in assembly order with
 * duplicates summed.  Any pointer may be NULL.  Returns nnz, or < 0. */
long long hifde_operator_csr(hifde_factor* f, int* rowptr, int* colind, double* vals);

/* Independent problems batched through shared launches (no reference
 * counterpart: the reference factors one problem at a time).
 * hifde_create_batch makes one factor object for nprob (1 to 8) problems of
 * the same grid (N interior points per side, leaf P, epsilon): they are
 * packed as the octants of a grid 2(N+1) points wide (problem o at offset
 * (N+1) * (o & 1, o >> 1 & 1, o >> 2 & 1)), the planes between the octants
 * and any unused octant carrying identity rows with no couplings.  Each
 * level of the factorization then eliminates the boxes of every problem in
 * the same launches; every box of a problem is its standalone box (the
 * separating points are extra, decoupled DOFs), except that the FRONT-face
 * guard of src/Factor.cpp:486 applies at the packed grid's k = 0 only.
 * hifde_set_stencil_batch_device assembles problem o's stencil from its
 * device nodal tensors d_a[o], d_v[o] (host arrays of nprob device
 * pointers; tensors as hifde_set_stencil_device).  hifde_batch_pack moves
 * vectors between the problems' own layout -- vector c of problem o at
 * d_sub + (o*nrhs + c)*ld_sub, (N+1)^3 entries -- and the packed layout --
 * column c at d_packed + c*ld_packed, (2(N+1))^3 entries -- (unpack != 0:
 * packed to problems).  hifde_pcg_multi_device on a batch runs one PCG per
 * (problem, column) with its own dots, step lengths and stopping test;
 * iters and relres then hold nprob*nrhs entries, [o*nrhs + c]. */
int hifde_create_batch(int N, int P, double epsilon, int precision, int flags, int nprob,
                       hifde_factor** out);
int hifde_set_stencil_batch_device(hifde_factor* f, const double* const* d_a,
                                   const double* const* d_v, int symmetric);
int hifde_batch_pack(hifde_factor* f, double* d_sub, long long ld_sub, double* d_packed,
                     long long ld_packed, int nrhs, int unpack);

/* Synthetic-input generators on device memory (no reference counterpart:
 * the reference fills its tensors by hand with a = 1, V = 0,
 * tests/HIFDE3D/main.cpp:59-74).  Each runs on the given stream (a
 * hipStream_t as void*, NULL for the legacy default stream).  Number t of a
 * seed's sequence is the SplitMix64 output t + 1 (seed + (t+1) * 0x9E3779B97F4A7C15
 * through the published finaliser); a uniform is its top 53 bits / 2^53.
 *   hifde_gen_uniform: d_out[t] = lo + (hi - lo) u_t, or 10^(that) with pow10
 *   hifde_gen_normal: Box-Muller, d_out[2p] = r cos(2 pi u_{2p+1}),
 *     d_out[2p+1] = r sin(2 pi u_{2p+1}), r = sqrt(-2 log(1 - u_{2p}))
 *   hifde_gen_rhs_normal: nrhs vectors of NC^3 (column c at d_out + c*ld):
 *     normal numbers c*N^3 + rank on the interior DOFs (rank = position in
 *     linear-index order), 0 on the boundary planes
 *   hifde_gen_gaussian_field: unit-variance Gaussian random field on the
 *     (NC+1)^3 nodes of the unit cube with covariance exp(-|x-y|^2/(2 corr^2)):
 *     normal white noise (numbers in order) on the doubled grid
 *     P = 2(NC+1), filtered by the square root of the spectral density
 *     (a periodic separable filter on spacing 1/NC), the first NC+1 nodes per
 *     axis kept, shifted to mean 0 and scaled to unit population variance;
 *     with lognormal != 0, exp of it.  Returns after the stream completes. */
int hifde_gen_uniform(double* d_out, long long n, unsigned long long seed, double lo, double hi,
                      int pow10, void* stream);
int hifde_gen_normal(double* d_out, long long n, unsigned long long seed, void* stream);
int hifde_gen_rhs_normal(double* d_out, int NC, int nrhs, long long ld, unsigned long long seed,
                         void* stream);
int hifde_gen_gaussian_field(double* d_out, int NC, unsigned long long seed, double corr,
                             int lognormal, void* stream);

/* Initialize() + Factor() -- src/Factor.cpp:6-39, 41-63.  The level matrix is
 * consumed, as in the reference; the assembled operator is kept for SpMV
 * and PCG. */
int hifde_factor_run(hifde_factor* f);

/* Apply(u, apply_inverse) -- src/Factor.cpp:992-1081.  u: host vector of
 * (N+1)^3 doubles, overwritten with F u or F^{-1} u (boundary entries are
 * zeroed first, src/Factor.cpp:1001). */
int hifde_apply(hifde_factor* f, double* u, int inverse);
/* Same on device memory: nrhs vectors at d_u + c*ldu.  For nrhs > 1 every
 * box step is one batched GEMM over the nrhs columns (the factor blocks are
 * read once per sweep); mixed-precision levels then round the vectors to
 * FP32 for the product, as their factor blocks are FP32.
 * The device entry points (this one and the PCG ones) run on the factor's
 * own stream, after the work already queued on the caller's stream (see
 * hifde_set_stream; without blocking the host), and return when done;
 * buffers written on any other stream must be complete before the call.
 * Repeated calls with the same (d_u, nrhs, ldu, inverse) replay the sweeps
 * from a HIP graph (option "graphs"); if HIP refuses or invalidates the
 * capture (for instance because another host thread synchronizes the whole
 * device meanwhile), the call runs the same sweeps eagerly, with the same
 * results, and the object stops capturing. */
int hifde_apply_device(hifde_factor* f, double* d_u, int nrhs, long long ldu, int inverse);
/* hifde_apply_device for one vector (nrhs = 1) that allocates nothing: no
 * hipMalloc, hipFree, event, graph or other allocating call.  No counterpart
 * in the reference; it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * Its scratch (3 (N+1)^3 doubles) is a fixed array in the library, one copy
 * per GPU, resolved when a factor object is created; (N+1)^3 above 128^3
 * returns HIFDE_ERR_SIZE.  Same kernels, same order and same results as
 * hifde_apply_device with nrhs = 1, run eagerly (never from a graph) and
 * without kernel timing.  Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (from several threads or objects) race. */
int hifde_apply_device_noalloc(hifde_factor* f, double* d_u, int inverse);
/* The stream (a hipStream_t, passed as void*) on which the caller produces
 * the device inputs of the *_device entry points and consumes their outputs;
 * NULL (the default) is the legacy default stream.  No reference
 * counterpart (the reference is host-only).  A multithreaded caller names a
 * stream of its own per thread, so that no call touches the legacy stream. */
int hifde_set_stream(hifde_factor* f, void* stream);

/* SpMV(A, x, y) -- src/Factor.cpp:1099-1113.  transpose == 1 computes the
 * reference's y = A^T x; transpose == 0 computes y = A x.  Host vectors. */
int hifde_spmv(hifde_factor* f, const double* x, double* y, int transpose);

/* Preconditioned CG on A x = b with F^{-1} as preconditioner (no reference
 * counterpart; the reference stops at Apply).  FP64 residuals and dots.
 * Stops when ||b - A x|| <= tol ||b||.  x holds the initial guess on entry. */
int hifde_pcg(hifde_factor* f, const double* b, double* x, double tol, int maxit, int* iters,
              double* relres);
int hifde_pcg_device(hifde_factor* f, const double* d_b, double* d_x, double tol, int maxit,
                     int* iters, double* relres);
/* nrhs independent PCG solves advanced together (no reference counterpart):
 * column c of B at d_b + c*ldb, of X at d_x + c*ldx (initial guesses on
 * entry).  Every column follows hifde_pcg's recurrence and stopping rule; a
 * column that has converged is left unchanged while the others continue.
 * iters[c] and relres[c] receive each column's count and final residual.
 * The preconditioner applies F^{-1} to all columns at once (batched GEMMs
 * per box), as hifde_apply_device does for nrhs > 1. */
int hifde_pcg_multi_device(hifde_factor* f, const double* d_b, long long ldb, double* d_x,
                           long long ldx, int nrhs, double tol, int maxit, int* iters,
                           double* relres);

/* Factor persistence (no reference counterpart; the reference keeps its
 * factor in memory only).  hifde_factor_save writes the operator, every box's
 * DOF lists and its factor blocks (A22, A22^{-1}, X, W in the level's
 * precision) to a binary file; hifde_factor_load reads it into an object
 * created with the same N, P, epsilon, precision and flags (HIFDE_ERR_ARG
 * otherwise), which can then apply, solve and re-factor as if it had factored
 * itself.  Returns the saved factorization's status (>= 0) on success.  A
 * load that fails part-way (truncated or corrupt file) leaves the object
 * without a factorization; a header mismatch leaves it untouched. */
int hifde_factor_save(hifde_factor* f, const char* path);
int hifde_factor_load(hifde_factor* f, const char* path);

/* Introspection (the reference prints these, src/Factor.cpp:211-212,258-259). */
int hifde_num_levels(const hifde_factor* f);
/* out[0..4] = cells, DOFs eliminated by cells, faces, DOFs eliminated by
 * faces, skeleton DOFs kept by faces */
int hifde_level_counts(const hifde_factor* f, int level, int* out);
/* kind 0 = cells (schur_level_data), 1 = faces (skel_level_data) */
int hifde_box_count(const hifde_factor* f, int level, int kind);
/* out[0..1] = |red|, |skel| */
int hifde_box_sizes(const hifde_factor* f, int level, int kind, int box, int* out);
/* FactorData of one box (include/data.hpp:83-110): global ids of red and
 * skel DOFs in the box's order, A_22, A_22^{-1}, X (r x s) and, for faces,
 * W (s x r).  Any pointer may be null; A22 must be null when the factor
 * was made with option "keep_a22" = 0 (HIFDE_ERR_STATE otherwise). */
int hifde_box_data(const hifde_factor* f, int level, int kind, int box, int* red_gid,
                   int* skel_gid, double* A22, double* A22inv, double* X, double* W);
/* Wall-clock seconds of the last factor by stage; returns the number of
 * stages written (names from hifde_stage_name). */
int hifde_stage_times(const hifde_factor* f, double* out, int nout);
/* The same, for one level of the last factor. */
int hifde_level_stage_times(const hifde_factor* f, int level, double* out, int nout);
const char* hifde_stage_name(int stage);
/* Tuning options (no reference counterpart):
 *   "gram_min_n"   faces with more ID columns than this use the Gram-matrix
 *                  path (G = M^T M, pivoted Cholesky); default 16
 *   "gram_min_eps" below this ID tolerance every face uses the Householder
 *                  column-pivoted QR; default 1e-7
 *   "fp32_levels"  HIFDE_MIXED only: levels 0 .. fp32_levels-1 are factored
 *                  and stored in FP32, the levels above in FP64; default
 *                  num_levels (every level in FP32) when epsilon >= 1e-5,
 *                  else num_levels - 2 (the two top levels in FP64)
 *   "symmetric"    -1 (default): when the assembled operator equals its
 *                  transpose bitwise, form only the lower triangle of each
 *                  Schur complement S = -A21^T A22^{-1} A21 (and store the
 *                  hash-table level matrix by unordered pairs); 0: general
 *                  path
 *   "stage_sync"   1: drain the stream at every stage boundary so that the
 *                  stage times include device time; default 0 (stage times
 *                  are host time and the host runs ahead of the device)
 *   "elements"     -1 (default): when every coupling of the operator lies
 *                  inside one level-0 cell closure (all stencils), hold the
 *                  level matrix as per-cell dense elements assembled by
 *                  extend-add; 0: device hash table of (row, col) entries
 *   "elements_fp32"  -1 (default): when every level is factored in FP32
 *                  (HIFDE_MIXED with fp32_levels = num_levels) the elements
 *                  are FP32 as well, so their assembly, face gathers and
 *                  Schur-complement accumulation move half the bytes (the
 *                  face ID matrices are still gathered into FP64); 0: FP64
 *                  elements always
 *   "symmetric_inverse"  1 (default): with the symmetric level matrix the
 *                  block inverses read only lower triangles and update only
 *                  lower triangles (half the flops); 0: general inverse
 *   "bf16x3"       0 (default); 1: on the FP32 levels of HIFDE_MIXED, the
 *                  Schur products X = A22^{-1} A21 and S = -A12 X of depth
 *                  above 64 run as bf16x3 split products (see HIFDE_BF16X3);
 *                  inverses, IDs and the FP64 levels are unchanged
 *   "gemm_tile64"  1 (default): products of depth above 64 whose 128 x 128
 *                  matrix-core tiles would compute over a quarter more output
 *                  than 64 x 64 ones run on 64 x 64 tiles; 0: always
 *                  128 x 128; 2: also products of depth 64 or less with both
 *                  output sides of at least 32 (otherwise on the FMA kernel)
 *   "keep_workspaces"  1 (default): the factorization's workspaces (level-
 *                  matrix elements, face and Schur blocks, inverse panels)
 *                  stay allocated for the next hifde_factor_run; 0 frees
 *                  them after each factorization (and at once if factored),
 *                  leaving the device memory to the factor and the solves
 *                  (at NC=256, P=8 the workspaces exceed the factor)
 *   "keep_a22"     1 (default): every box keeps A22 beside A22^{-1} (as the
 *                  reference's FactorData); 0: each box is inverted where it
 *                  was gathered and the A22 blocks are freed when the
 *                  factorization ends (the factor shrinks by a third or
 *                  more); Apply F^{-1}, PCG and every solve give the same
 *                  bits, the forward Apply F returns HIFDE_ERR_STATE
 *   "plan_ahead"   1 (default): with the element storage the host plans the
 *                  level-1 cells while the device runs the level-0 face
 *                  IDs, and keeps that plan when those faces eliminate no
 *                  DOF; 0: planned after them.  Same results either way
 *   "gj_fused_steps"  1 (default): each inner step of the blocked inverses
 *                  launches its column kernel together with the next step's
 *                  diagonal block; 0: separate launches.  Same bits
 *   "gj_window"    0 (default: 512) or a multiple of 64 in [64, 4096]: the
 *                  outer window width of the blocked inverses (same results
 *                  to rounding)
 *   "host_threads" OpenMP threads of the host-side planning in
 *                  hifde_factor_run (0, the default: the calling thread's
 *                  setting); for several factor objects run from several
 *                  host threads at once
 *   "gram_panel_reg"  1 (default): Gram pivoted-Cholesky panels of faces up
 *                  to 1536 columns hold the panel rows in registers; 0: the
 *                  LDS/global-memory panel for every face (same results) */
int hifde_set_option(hifde_factor* f, const char* name, double value);
/* Per-kernel-family device time, recorded with HIP events on the factor's
 * stream while profiling is on.  hifde_kernel_stats returns the number of
 * families; for 0 <= idx < that number it fills name and
 * out[0..3] = launches, seconds, algorithmic flops, algorithmic bytes. */
int hifde_set_profiling(hifde_factor* f, int on);
/* While profiling, time only the launches of the kernel with this name (as
 * listed by hifde_kernel_stats without the "kernel:" prefix) and record no
 * family totals; NULL or "" restores full profiling.  Keeps the event
 * overhead of a timed run to that one kernel. */
int hifde_set_profile_kernel(hifde_factor* f, const char* kernel);
int hifde_reset_kernel_stats(hifde_factor* f);
int hifde_kernel_stats(const hifde_factor* f, int idx, char* name, int name_len, double* out);
/* Bytes of factor data held on the device. */
long long hifde_factor_bytes(const hifde_factor* f);

/* InterpDecomp(M, W, skel, red, eps) -- src/InterpDecomp.cpp:86-138, over a
 * batch.  M_packed: the m[b] x n[b] matrices packed one after another; as in
 * the reference (include/InterpDecomp.hpp:12, "M gets overwritten") each is
 * overwritten with its column-pivoted QR factors (see
 * hifde_pivoted_qr_batch).  Outputs per face: k_out[b]; perm_out (sum n[b]
 * ints): the k skeleton columns then the n-k redundant columns, in the
 * reference's order; W_out (sum n[b]^2 doubles, face b at offset
 * sum_{c<b} n[c]^2): W, k x (n-k). */
int hifde_interp_decomp_batch(int nbatch, const int* m, const int* n, double* M_packed,
                              double eps, int* k_out, int* perm_out, double* W_out);
/* PivotedQRWrapper(A, jpvt) -- src/InterpDecomp.cpp:156-166,198-211: LAPACK
 * dgeqp3 with every column free, pivots made 0-based; over a batch.  A_packed
 * (in/out): the m[b] x n[b] matrices packed one after another, overwritten
 * with R on and above the diagonal and the Householder vectors below it
 * (dgeqp3's layout, unit leading entries implicit).  jpvt_out (sum n[b]
 * ints): column i of A P is column jpvt[i] of A.  tau_out (may be NULL;
 * sum min(m[b], n[b]) doubles): the reflector scalars.  FP64 arithmetic on
 * the GPU (the factorization's ID kernel). */
int hifde_pivoted_qr_batch(int nbatch, const int* m, const int* n, double* A_packed,
                           int* jpvt_out, double* tau_out);
/* TriangularSolveWrapper(R, B) -- src/InterpDecomp.cpp:213-231 (dtrsm: left,
 * upper, no transpose, non-unit diagonal), over a batch: B_b := R_b^{-1} B_b
 * with R_b (m[b] x m[b], upper triangle read) packed one after another in
 * R_packed and B_b (m[b] x n[b]) likewise in B_packed (overwritten).  FP64
 * on the GPU; entries that are exactly zero are skipped as reference BLAS
 * dtrsm skips them, and a zero diagonal divides by zero as dtrsm does. */
int hifde_triangular_solve_batch(int nbatch, const int* m, const int* n, const double* R_packed,
                                 double* B_packed);
/* Schur(matrix, data) -- include/Schur.hpp:72-100, over a batch, given the
 * three blocks A22 = M(red,red) (r x r), A21 = M(red,skel) (r x s),
 * A12 = M(skel,red) (s x r), packed per box.  Outputs packed A22^{-1}, X and
 * S = -A12 X.  precision selects FP64 or FP32 arithmetic; OR-ed with
 * HIFDE_SCHUR_SYMMETRIC it declares every A22 symmetric, and the inverses of
 * blocks wider than 512 take half the flops (no counterpart in the
 * reference, whose Schur inverts by LU).  info[b] as LU. */
#define HIFDE_SCHUR_SYMMETRIC 0x100
/* OR-ed with HIFDE_MIXED into the precision of hifde_schur_batch (its X and S
 * products) or hifde_gemm_batch: matrix-core products of depth above 64 are
 * bf16x3 split products (each FP32 operand split into two bf16 parts, three
 * bf16 products summed in FP32; about 2^-16 relative per product instead of
 * FP32's 2^-24).  No counterpart in the reference, which multiplies in the
 * factor's precision. */
#define HIFDE_BF16X3 0x200
/* OR-ed into the precision of hifde_gemm_batch: every product with both
 * output sides of at least 32 runs on 64 x 64 matrix-core tiles (the kernel
 * behind the factorization's option "gemm_tile64").  No counterpart in the
 * reference. */
#define HIFDE_GEMM_TILE64 0x400
/* OR-ed into the precision of hifde_gemm_batch: run every product on one
 * kernel, the 64 x 64 FMA tiles (FMA), the 128 x 128 matrix-core tiles
 * (CORE128) or the 64 x 64 matrix-core tiles (CORE64); a product the
 * matrix-core kernel is not validated for (an output side below 32, depth 0)
 * stays on the FMA kernel.  The candidates of the GEMM tuner; no counterpart
 * in the reference. */
#define HIFDE_GEMM_FORCE_FMA 0x800
#define HIFDE_GEMM_FORCE_CORE128 0x1000
#define HIFDE_GEMM_FORCE_CORE64 0x2000
/* OR-ed into the precision of hifde_gemm_batch: route by the fixed rule (the
 * flags above), ignoring a loaded routing table (hifde_gemm_table_load), whose
 * entries otherwise take precedence over the rule for the classes they cover.
 * No counterpart in the reference. */
#define HIFDE_GEMM_RULE 0x4000
/* OR-ed into the precision of hifde_gemm_batch: square products (m == n)
 * compute only the output tiles on or below the block diagonal, as the
 * factorization's symmetric products (S = -A21^T X, the symmetric inverse's
 * trailing updates) do; only the lower triangle of such a C is defined.
 * Products with m != n are computed whole.  No counterpart in the
 * reference. */
#define HIFDE_GEMM_LOWER 0x8000
int hifde_schur_batch(int precision, int nbatch, const int* r, const int* s, const double* A22,
                      const double* A21, const double* A12, double* A22inv, double* X, double* S,
                      int* info);

/* hmat_tools::Multiply(alpha=1, A, B, C) -- src/hmat_tools/MultiplyMatrix.cpp:15-69
 * (C := A B) and AdjointMultiply (C := A^T B, AdjointMultiplyMatrix.cpp:15-96)
 * over a batch: C_b (m x n) = op(A_b) B_b, op(A) = A (m x k) or A^T (A stored
 * k x m) when transA != 0.  Operands packed one after another, column-major.
 * precision selects FP64 or FP32 arithmetic (the kernels the factorization's
 * Schur and Gram products use). */
int hifde_gemm_batch(int precision, int transA, int nbatch, const int* m, const int* n,
                     const int* k, const double* A, const double* B, double* C);

/* GEMM routing tuner (no reference counterpart; the reference calls BLAS).
 * Every batched product of the factorization is routed per descriptor to
 * one of three kernels -- 0: FMA 64 x 64 tiles, 1: matrix-core 128 x 128,
 * 2: matrix-core 64 x 64 -- by a process-wide table indexed by the
 * descriptor's class (element size, transA, lower, depth bin, bin of the
 * smaller output side, bin of the number of descriptors in the call), where
 * the table has an entry, and by the fixed rule otherwise.  The table is
 * written by hifde3d_amd/tune.py, which times the candidates on the
 * products the workloads record; the one shipped next to the library
 * (gemm_tune.txt) is loaded when the first factor object is created.
 *   hifde_gemm_table_load: load a table file (NULL: the shipped one);
 *     returns the number of entries.  hifde_gemm_table_clear: fixed rule
 *     everywhere.  Option "gemm_table" = 0 makes one object ignore it.
 *   hifde_gemm_log: with option "gemm_record" = 1, the descriptors of the
 *     last factorization, 9 ints each (call, esize, transA, lower, m, n, k,
 *     descriptors in the call, kind routed); returns the int count (copies
 *     at most cap).
 *   hifde_gemm_time: average device seconds over reps launches of a batch
 *     of packed products (m[b] x n[b], depth k[b]; lower: lower-triangle
 *     tiles only) forced onto one kind, operands generated on the GPU. */
int hifde_gemm_table_load(const char* path);
int hifde_gemm_table_clear(void);
long long hifde_gemm_log(const hifde_factor* f, int* out, long long cap);
int hifde_gemm_time(int precision, int transA, int lower, int nbatch, const int* m, const int* n,
                    const int* k, int kind, int reps, double* avg_seconds);

#ifdef __cplusplus
}
#endif

#endif /* HIFDE_AMD_H_ */
