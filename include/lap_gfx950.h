/*
 * lap_gfx950.h -- C ABI of the MI355X (gfx950) backend for the LAP PCG solver.
 *
 * liblapgfx.so exports three groups of symbols:
 *
 *  1. The reference's HIP backend surface (hip_blas.h + devMem.h), symbol for
 *     symbol, so the reference's host code (cg.c, prec.c, GS.c, it_jacobi.c,
 *     line_jacobi.c, blas.c, io_utils.c, cg_driver*.c, mm_driver.c built with
 *     -DHIP=1) links against this library in place of hip_blas.o + devMem.o +
 *     rocSPARSE + rocBLAS.
 *  2. The reference's solver layer (common.h:52-74, blas.h:3-48): cg(),
 *     prec_function(), the preconditioners and the blas.h wrappers.  Linking
 *     these from the library instead of the reference's objects switches the
 *     solve to the fused, device-resident implementation.  A definition in the
 *     executable takes precedence over the library's (ELF interposition), so
 *     both link modes use the same library.
 *  3. lap_* extensions with no reference counterpart (status reporting, the
 *     explicit solve entry points, mixed precision, block and batched PCG,
 *     adjoint gradients).
 *
 * Conventions (as the reference): every pointer argument of groups 1-2 is a
 * device pointer allocated with mallocForDevice unless stated otherwise;
 * scalars alpha/beta of hip_csr_matvec are host pointers; hip_dot returns a
 * host value; every call is synchronous at return; the library keeps
 * per-process state (analysis plans) and is not thread-safe.  Errors are
 * printed to stderr and recorded; lap_last_status()/lap_last_error() return
 * them (the reference's functions return void).
 */
#ifndef LAP_GFX950_H
#define LAP_GFX950_H

#ifdef __cplusplus
extern "C" {
#endif

/* Preconditioner state.  Same member order and types as the reference's
 * pdata (common.h:29-50, with real_type = double), so a pdata* from the
 * reference's drivers can be passed where a lap_pdata* is expected. */
typedef struct lap_pdata {
  int *lia;
  int *lja;
  double *la;
  int lnnz;
  int *uia;
  int *uja;
  double *ua;
  int unnz;
  double *ichol_vals;
  double *d;
  double *d_r;
  int n;
  double *aux_vec1, *aux_vec2, *aux_vec3;
  char *prec_op;
  int m, k;
} lap_pdata;

/* ===================== 1. hip_blas.h / devMem.h surface ================== */

/* hip_blas.h:5 (hip_blas.cpp:23-44): creates the library's stream and
 * device workspace.  Optional: every entry point initialises lazily. */
void initialize_handles(void);

/* hip_blas.h:7-14 (hip_blas.cpp:46-100).  Builds the SpMV plan (row blocks
 * binned by row length) for the CSR arrays (ia, ja); option "A", "L" or "U"
 * is accepted for compatibility.  Plans are keyed by (ia, ja, n, nnz), so
 * hip_csr_matvec also works on matrices that were never analysed. */
void analyze_spmv(const int n, const int nnz, int *ia, int *ja, double *a,
                  const double *x, double *result, char *option);

/* hip_blas.h:16-24 (hip_blas.cpp:102-173).  Level analysis of L+D (diagonal
 * last in each row) and D+U (diagonal first) for the triangular solves. */
void initialize_and_analyze_L_and_U_solve(const int n, const int nnzL, int *lia, int *lja,
                                          double *la, const int nnzU, int *uia, int *uja,
                                          double *ua);

/* hip_blas.h:26 (hip_blas.cpp:432-445): returns v.w on the host. */
double hip_dot(const int n, const double *v, const double *w);
/* hip_blas.h:28 (hip_blas.cpp:447-452): v *= alpha. */
void hip_scal(const int n, const double alpha, double *v);
/* hip_blas.h:30 (hip_blas.cpp:454-465): y += alpha x. */
void hip_axpy(const int n, const double alpha, const double *x, double *y);

/* hip_blas.h:32-41 (hip_blas.cpp:467-532): result = al*A*x + bet*result,
 * al/bet host pointers.  Rows are summed in CSR order.  bet == 0 does not
 * read result (BLAS convention; the CPU path computes result*0). */
void hip_csr_matvec(const int n, const int nnz, const int *ia, const int *ja, const double *a,
                    const double *x, double *result, const double *al, const double *bet,
                    const char *kind);

/* hip_blas.h:43-59 (hip_blas.cpp:534-591): result = T^{-1} x for a lower
 * (upper) triangular CSR T.  The pivot of row i is the entry with column i
 * when the row holds one (the GPU driver's L+D / D+U), else diagonal[i]
 * (the CPU path's strict L / U); off-diagonal terms are subtracted in CSR
 * order, as simple_blas.c:46-85. */
void hip_lower_triangular_solve(const int n, const int nnzL, const int *lia, const int *lja,
                                const double *la, const double *diagonal, const double *x,
                                double *result);
void hip_upper_triangular_solve(const int n, const int nnzU, const int *uia, const int *uja,
                                const double *ua, const double *diagonal, const double *x,
                                double *result);

/* hip_blas.h:61-71 (hip_blas.cpp:368-430, :597-630): element-wise ops.
 * reciprocal gives 0 where v == 0; sqrt gives 0 where v <= 0. */
void hip_vec_vec(const int n, const double *x, const double *y, double *res);
void hip_vector_reciprocal(const int n, const double *v, double *res);
void hip_vector_sqrt(const int n, const double *v, double *res);
void hip_vec_copy(const int n, const double *src, double *dest);
void hip_vec_zero(const int n, double *vec);
void hip_vec_set(const int n, double value, double *vec);

/* hip_blas.h:73-77 (hip_blas.cpp:175-317): IC(0) factorisation of the SPD
 * matrix (ia, ja, a) on the device, level-scheduled, written in place into
 * the lower triangle (diagonal included) of a; the upper triangle of a is
 * left unchanged.  The factor equals the reference CPU initialize_ichol's
 * (simple_blas.c:139-182) bitwise. */
void initialize_ichol(const int n, const int nnzA, int *ia, int *ja, double *a);

/* hip_blas.h:79-85 (hip_blas.cpp:319-366): y = L^{-T} L^{-1} x with the
 * factor set up by initialize_ichol on prec_data->ichol_vals; the
 * arithmetic order is simple_ichol's (simple_blas.c:184-223). */
void hip_ichol(const int *ia, const int *ja, double *a, const int nnzA, lap_pdata *prec_data,
               double *x, double *y);

/* devMem.h:2-4 (devMem.cpp:9-48).  memcpyDevice's type is "H2D", "D2H" or
 * "D2D" (D2D copies device to device; the reference passes
 * DeviceToHost there, devMem.cpp:33-39). */
void *mallocForDevice(void *x, int n, int size);
void memcpyDevice(void *dest, void *src, int n, int size, char *type);
void freeDevice(void *p);

/* ===================== 2. solver layer (common.h, blas.h) ================ */

/* common.h:54-66 (cg.c:5-140): PCG from x0 = x.  Device-resident: the
 * preconditioner, SpMV, dots and updates stay on the GPU, alpha/beta and
 * the convergence test are computed on the device, iterations are replayed
 * from a HIP graph and the host polls for convergence.  The per-iteration
 * "CG: it %d, res norm %5.5e" lines and the "TRUE Norm of r" line are
 * printed after the solve from the device residual history.
 * res_norm_history is a host array of at least maxit+2 entries; *flag is
 * 0 converged, 1 maxit reached, 2 failure (non-finite value or breakdown).
 * Recognised prec_op: none, line_jacobi, it_jacobi, GS_std, GS_it, GS_it2,
 * ichol. */
void cg(int n, double nnz, int *ia, int *ja, double *a, double *x, double *b, double tol,
        lap_pdata *prec_data, int maxit, int *it, int *flag, double *res_norm_history);

/* common.h:52 (prec.c:6-45) and common.h:70-74 */
void prec_function(int *ia, int *ja, double *a, int nnzA, lap_pdata *prec_data, double *x,
                   double *y);
void GS_std(int *ia, int *ja, double *a, int nnzA, lap_pdata *prec_data, double *vec_in,
            double *vec_out);
void GS_it(int *ia, int *ja, double *a, int nnzA, lap_pdata *prec_data, double *vec_in,
           double *vec_out);
void GS_it2(int *ia, int *ja, double *a, int nnzA, lap_pdata *prec_data, double *vec_in,
            double *vec_out);
void it_jacobi(int *ia, int *ja, double *a, int nnzA, lap_pdata *prec_data, double *vec_in,
               double *vec_out);
void line_jacobi(int *ia, int *ja, double *a, int nnzA, lap_pdata *prec_data, double *vec_in,
                 double *vec_out);

/* blas.h:3-48 (blas.c:15-186 with -DHIP=1): forward to the hip_* above. */
double dot(const int n, const double *v, const double *w);
void axpy(const int n, const double alpha, double *x, double *y);
void scal(const int n, const double alpha, double *v);
void csr_matvec(const int n, const int nnz, const int *ia, const int *ja, const double *a,
                const double *x, double *result, const double *al, const double *bet,
                const char *kind);
void lower_triangular_solve(const int n, const int nnz, const int *lia, const int *lja,
                            const double *la, const double *diag, const double *x,
                            double *result);
void upper_triangular_solve(const int n, const int nnz, const int *uia, const int *uja,
                            const double *ua, const double *diag, const double *x,
                            double *result);
void ichol(const int *ia, const int *ja, double *a, const int nnzA, lap_pdata *prec_data,
           double *x, double *y);
void vec_vec(const int n, const double *x, double *y, double *res);
void vector_reciprocal(const int n, const double *v, double *res);
void vector_sqrt(const int n, const double *v, double *res);
void vec_copy(const int n, double *src, double *dest);
void vec_zero(const int n, double *vec);

/* ===================== 3. lap_* extensions =============================== */
/* No reference counterpart. */

#define LAP_OK 0
#define LAP_ERR_ARG 1      /* invalid argument or sizes                   */
#define LAP_ERR_HIP 2      /* a HIP runtime call failed                   */
#define LAP_ERR_STRUCT 3   /* matrix structure not as required            */
#define LAP_ERR_NUMERIC 4  /* zero/negative pivot, non-finite value        */
#define LAP_ERR_PREC 5     /* unknown preconditioner name                 */

int lap_last_status(void);
const char *lap_last_error(void);
void lap_clear_error(void);

/* Drops every cached analysis plan, factor and work buffer. */
void lap_release_all(void);
/* Plans (SpMV row blocks, triangular level sets, IC0 factors, the mixed-
 * precision copy, batch orders) are cached by the device addresses of the
 * arrays they were built from.  Call this when an address is about to hold a
 * different matrix (e.g. after freeing and reallocating): every plan whose
 * key holds p is dropped.  analyze_spmv and
 * initialize_and_analyze_L_and_U_solve always rebuild their plans. */
void lap_forget(const void *p);

/* Result of one PCG solve. */
typedef struct lap_solve_info {
  int iterations;    /* as *it of cg()                                     */
  int flag;          /* 0 converged, 1 maxit, 2 failure                     */
  double res0;       /* ||b - A x0||                                        */
  double res_final;  /* recursive residual norm at exit                     */
  double true_res;   /* ||b - A x|| recomputed at exit                      */
  double gpu_ms;     /* device time of the iteration loop (HIP events)      */
  int refinements;   /* lap_pcg_mixed: FP64 residual replacements            */
} lap_solve_info;

/* cg() without printing: the same solve, explicit result.  history (host)
 * may be NULL or hold maxit+2 entries.  Returns a LAP_* status. */
int lap_pcg(int n, int nnz, const int *ia, const int *ja, const double *a, double *x,
            const double *b, double tol, lap_pdata *prec_data, int maxit, double *history,
            lap_solve_info *info);

/* Mixed-precision PCG for the same system: an FP32 inner PCG (FP32 vectors,
 * FP32 off-diagonal values of A, FP32 preconditioner) inside FP64 iterative
 * refinement with reliable updates: whenever the FP32 recursive residual has
 * dropped by `delta` (0 selects 0.01) since the last refinement, the FP32
 * correction is added to the FP64 x and the residual is replaced by the FP64
 * b - A x, without restarting the Krylov process.  FP64 is kept for x, b,
 * the refinement residual (the caller's FP64 CSR), the diagonal of A, every
 * row sum and every dot product.  Convergence (||b - A x|| < tol ||r0||) is
 * decided on FP64 residuals only.  prec_op: line_jacobi, ichol (factor from
 * initialize_ichol on ichol_vals) or none.  history gets the residual norm
 * per inner iteration (FP64 at refinements, FP32 recursive otherwise). */
int lap_pcg_mixed(int n, int nnz, const int *ia, const int *ja, const double *a, double *x,
                  const double *b, double tol, lap_pdata *prec_data, int maxit, double delta,
                  double *history, lap_solve_info *info);

/* Batched Jacobi PCG: nsys independent SPD systems, concatenated into one
 * block-diagonal CSR (global row/column indices; system s owns rows
 * [row_off[s], row_off[s+1])) with concatenated x (initial guess in, solution
 * out), b and d_r = 1/diag.  One persistent launch: workgroups take whole
 * systems from an atomic queue (largest first) and run each system's PCG
 * (cg.c recurrence, line_jacobi) with workgroup-level reductions only.
 * block (256/512/1024) and tpr (threads per row: 1/2/4) select the kernel;
 * block = -1 selects the lockstep mode instead (every system iterates
 * together over the block-diagonal matrix, several launches per iteration
 * from a HIP graph, stopped systems skipped: for batches whose largest
 * system would dominate one workgroup); 0/0 uses the configuration found by
 * lap_pcg_batched_tune (or 256/1).
 * Per-system outputs (host arrays of nsys, each may be NULL): iterations,
 * flag (0/1/2 as cg()), ||r0|| and the final recursive residual. */
int lap_pcg_batched(int nsys, const int *row_off, int n_total, int nnz_total, const int *ia,
                    const int *ja, const double *a, const double *d_r, double *x,
                    const double *b, double tol, int maxit, int block, int tpr, int *iters,
                    int *flags, double *res0, double *res, double *gpu_ms);
/* Autotuner for lap_pcg_batched: times trial_iters iterations of every
 * (block, tpr) kernel and of the lockstep mode on this batch (x = 0, the
 * caller's x is not touched) and keeps the fastest for later calls with
 * block = 0; *best_block = -1 when the lockstep mode won. */
int lap_pcg_batched_tune(int nsys, const int *row_off, int n_total, int nnz_total,
                         const int *ia, const int *ja, const double *a, const double *d_r,
                         const double *b, int trial_iters, int *best_block, int *best_tpr);

/* Multi-RHS (no reference counterpart; the reference's cg_driver2.c solves
 * one right-hand side per run): k independent line_jacobi PCG solves of one
 * matrix, each with cg.c's recurrence and stopping rule, sharing one SpMM
 * per iteration so the matrix is read once for all k columns.
 * X and B are n x k row-interleaved (element (i, j) at i*k + j), device
 * memory; k is a power of two <= 64; d_r = 1/diag(A) as pdata.d_r.
 * tile: rows per lane group per step of the SpMM (1, 2 or 4; one lane per
 * column, 64/k row groups per wave; with k = 64 also 8 or 16: rows per wave
 * step, indices and values loaded by the whole wave and broadcast), 0 = pick
 * by timing (kept per matrix and k).  hist (host, may be NULL):
 * (maxit+2) x k residual norms, row it = iteration it (0 past a column's
 * own stop).  Per-column outputs
 * (host arrays of k, each may be NULL) as lap_pcg_batched. */
int lap_pcg_block(int n, int nnz, const int *ia, const int *ja, const double *a,
                  const double *d_r, int k, double *X, const double *B, double tol, int maxit,
                  int tile, double *hist, int *iters, int *flags, double *res0, double *res,
                  double *gpu_ms);
/* Y = A X for k row-interleaved columns (the SpMM of lap_pcg_block); each
 * column equals hip_csr_matvec's result bit for bit. */
int lap_csr_matmat(int n, int nnz, const int *ia, const int *ja, const double *a, int k,
                   const double *X, double *Y, int tile);

/* Differentiation (no reference counterpart; the reference builds an
 * unweighted Laplacian once on the host, io_utils.c:80-130, :364-420).
 * L(w) = sum_e w_e (u_s - u_d)(u_s - u_d)^T + shift*I over the undirected
 * edges e = (src[e], dst[e]) (no self loops, no duplicates), all device
 * arrays.
 * lap_graph_pattern: CSR pattern of L (n + 2m entries, rows sorted, the
 * diagonal always present) with epos[2e], epos[2e+1] the positions of
 * (s, d) and (d, s) and dpos[i] that of (i, i).  Host-side setup.
 * lap_graph_values: a = values of L(w) on that pattern (off-diagonals -w_e,
 * diagonal = the row's weights summed in CSR order, plus shift), and
 * d_r = 1/diag when d_r is not NULL (the line_jacobi operand).
 * lap_edge_grad: grad_w[e] = -(lam[s] - lam[d]) (x[s] - x[d]).
 * lap_pcg_adjoint: lam = L^{-1} grad_x by lap_pcg from lam = 0 with the
 * forward solve's preconditioner (prec_data), then lap_edge_grad when
 * grad_w is not NULL; dl/db = lam.  Returns LAP_ERR_NUMERIC if the adjoint
 * solve does not converge. */
int lap_graph_pattern(int n, int m, const int *src, const int *dst, int *ia, int *ja, int *epos,
                      int *dpos);
int lap_graph_values(int n, int m, const int *ia, const int *ja, const int *epos,
                     const int *dpos, const double *w, double shift, double *a, double *d_r);
int lap_edge_grad(int n, int m, const int *src, const int *dst, const double *lam,
                  const double *x, double *grad_w);
int lap_pcg_adjoint(int n, int nnz, const int *ia, const int *ja, const double *a,
                    lap_pdata *prec_data, const double *grad_x, double *lam, double tol,
                    int maxit, int m, const int *src, const int *dst, const double *x,
                    double *grad_w, double *history, lap_solve_info *info);

/* Kernel-level entry points used by benchmarks and tests (device pointers):
 * y = A x with the fused SpMV, launched `reps` times back to back on the
 * library stream; returns the average device milliseconds per launch. */
double lap_spmv_timed(int n, int nnz, const int *ia, const int *ja, const double *a,
                      const double *x, double *y, int reps);
/* Same for the PCG's fused kernel q = A p with p.q and alpha in its
 * epilogue (the dominant kernel of a Jacobi PCG iteration). */
double lap_pcg_spmv_timed(int n, int nnz, const int *ia, const int *ja, const double *a,
                          const double *p, double *q, int reps);

/* Library configuration knobs (for tuning sweeps): name/value pairs.
 * "spmv_rows" (rows per SpMV block), "spmv_walk" (row-block order of the
 * persistent SpMV grid: 0 cyclic, 3 groups of four consecutive blocks per
 * XCD), "spmv_wg_per_cu" (its workgroups per CU), "vec_grid" (workgroups
 * of the vector kernels, 1..4096), "graph_iters" (iterations per HIP graph
 * launch, 1..4096), "use_graph" (0 or 1), "spmm_wg_per_cu" (block CG SpMM
 * workgroups per CU), "sptrsv_mode" (triangular solves: 0 one launch per
 * level, 3, the default, one persistent launch over chunks of rows with
 * point-to-point waits), "sptrsv_long" (1: rows of more than 32 entries
 * summed in a fixed tree order, rows of more than 1024 entries over several
 * workgroups), "fuse_sweeps", "spmv_format" (0 CSR, 1 value-indexed,
 * 2 delta-indexed), "mixed_vectors" (lap_pcg_mixed inner vectors: 0 FP32,
 * 1 FP64).  Results do not depend on them except vec_grid, spmv_walk and
 * spmv_wg_per_cu (the dot-product summation order), sptrsv_long and
 * mixed_vectors.  Returns the previous value, or -1 for an unknown name. */
int lap_set_option(const char *name, int value);

/* No counterpart in the reference: hip_dot for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  *result (device memory) = v.w with
 * hip_dot's kernels, so with its bits; n >= 0, any size (its partial sums
 * never exceed 2048).  The partial sums live in a static 16 KiB device array,
 * one per GPU, looked up at initialize_handles or at the first call on a
 * device; no call allocates.  Calls that overlap on one GPU (other streams
 * or processes) race on that array.  Returns LAP_OK or an error code. */
int lap_dot_noalloc(const int n, const double *v, const double *w, double *result);

#ifdef __cplusplus
}
#endif
#endif
