/*
 * lap_oracle.h -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU restatement of the reference LAP solver's serial path, used only as the
 * checker by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg.  Nothing in lap_amd/ links, loads or calls this library.
 *
 * Every function follows one reference function operation for operation
 * (same loop order, same association of floating-point operations), so that
 * built with -ffp-contract=off it is bitwise equal to the reference's
 * simple_blas build.  The citation next to each declaration names the
 * reference lines it restates.  Pinning status of each function is listed in
 * lap_oracle.c's header.
 */
#ifndef LAP_ORACLE_H
#define LAP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- BLAS-1 (simple_blas.c:4-22) ------------------------------------- */
double orc_dot(int n, const double *v, const double *w);            /* simple_blas.c:4-10  */
void   orc_scal(int n, double alpha, double *v);                    /* simple_blas.c:12-16 */
void   orc_axpy(int n, double alpha, const double *x, double *y);   /* simple_blas.c:18-22 */

/* ---- sparse kernels ---------------------------------------------------- */
/* y = alpha*A*x + beta*y, row-sequential (simple_blas.c:24-44) */
void orc_csr_matvec(int n, const int *ia, const int *ja, const double *a,
                    const double *x, double *y, double alpha, double beta);

/* Triangular solves (simple_blas.c:46-85).  The diagonal of row i is the
 * entry whose column equals i when the row holds one (the GPU driver's
 * augmented L+D / D+U, cg_driver2.c:95-144, and the IC0 factors,
 * simple_blas.c:184-223); otherwise it is diag[i] (the strict L/U of the CPU
 * path).  Off-diagonal terms are subtracted in CSR order. */
void orc_lower_solve(int n, const int *ia, const int *ja, const double *a,
                     const double *diag, const double *x, double *y);
void orc_upper_solve(int n, const int *ia, const int *ja, const double *a,
                     const double *diag, const double *x, double *y);

/* element-wise (simple_blas.c:91-136) */
void orc_vec_vec(int n, const double *x, const double *y, double *res);
void orc_reciprocal(int n, const double *v, double *res);
void orc_sqrt(int n, const double *v, double *res);

/* IC0 (simple_blas.c:139-182): right-looking factorisation in place on the
 * upper factor U = L^T (rows sorted, diagonal first), then the transpose is
 * scattered into la (L pattern, rows sorted, diagonal last). */
void orc_ichol_factor(int n, const int *uia, const int *uja, double *ua,
                      const int *lia, const int *lja, double *la);
/* IC0 apply y = U^{-1} L^{-1} x (simple_blas.c:184-223); aux has n entries */
void orc_ichol_apply(int n, const int *lia, const int *lja, const double *la,
                     const int *uia, const int *uja, const double *ua,
                     const double *x, double *y, double *aux);

/* ---- preconditioners + PCG -------------------------------------------- */
enum orc_prec_kind {
  ORC_PREC_NONE = 0,
  ORC_PREC_LINE_JACOBI = 1,   /* line_jacobi.c:3-7 */
  ORC_PREC_IT_JACOBI = 2,     /* it_jacobi.c:4-32  */
  ORC_PREC_GS_STD = 3,        /* GS.c:5-42         */
  ORC_PREC_GS_IT = 4,         /* GS.c:45-92        */
  ORC_PREC_GS_IT2 = 5,        /* GS.c:95-159       */
  ORC_PREC_ICHOL = 6          /* simple_blas.c:184-223 */
};

/* Preconditioner operands, the oracle's counterpart of the reference's
 * pdata (common.h:29-50). */
typedef struct {
  int kind;
  int n, m, k;
  const int *lia, *lja; const double *la; /* strict L (CPU path) or L factor */
  const int *uia, *uja; const double *ua; /* strict U (CPU path) or U factor */
  const double *d, *d_r;                  /* diag(A) and 1/diag(A)          */
  double *aux1, *aux2, *aux3;             /* n-vectors of scratch            */
} orc_prec;

void orc_prec_apply(const int *ia, const int *ja, const double *a,
                    const orc_prec *p, const double *x, double *y);

/* PCG (cg.c:5-140, CPU branch).  hist needs maxit+2 entries.
 * quiet != 0 suppresses the per-iteration log lines. */
void orc_cg(int n, const int *ia, const int *ja, const double *a,
            double *x, const double *b, double tol, const orc_prec *p,
            int maxit, int *it, int *flag, double *hist, int quiet);

/* ---- OpenMP-structured variants (openmp_blas.c:21-52, :166-176) ------- */
/* Timed CPU baseline only: SpMV row-parallel on spmv_threads threads (the
 * reference hard-codes 16, openmp_blas.c:37), dot a parallel reduction on
 * dot_threads threads, every other vector op sequential.  Numerically these
 * differ from orc_cg only through the order of the dot-product sums. */
void orc_omp_cg(int n, const int *ia, const int *ja, const double *a,
                double *x, const double *b, double tol, const orc_prec *p,
                int maxit, int *it, int *flag, double *hist,
                int spmv_threads, int dot_threads);
void orc_omp_csr_matvec(int n, const int *ia, const int *ja, const double *a,
                        const double *x, double *y, double alpha, double beta,
                        int threads);

/* ---- weighted graph Laplacian and its adjoint gradient --------------- */
/* No reference counterpart (the reference builds an unweighted Laplacian,
 * io_utils.c:80-130, :364-420); restated from the definition
 * L(w) = sum_e w_e (u_s - u_d)(u_s - u_d)^T + shift*I.  On a CSR pattern
 * with epos/dpos maps: a[epos[2e]] = a[epos[2e+1]] = -w[e]; the diagonal of
 * row i = (sum of -a over the row's off-diagonal entries, CSR order) +
 * shift. */
void orc_graph_values(int n, int m, const int *ia, const int *ja, const int *epos,
                      const int *dpos, const double *w, double shift, double *a);
/* g[e] = -(lam[s] - lam[d]) * (x[s] - x[d]): dl/dw_e for L(w) x = b with
 * adjoint L(w) lam = dl/dx. */
void orc_edge_grad(int m, const int *src, const int *dst, const double *lam, const double *x,
                   double *g);

/* ---- input path (io_utils.c) ------------------------------------------ */
/* Matrix Market coordinate reader, serial (io_utils.c:15-74, adjacency mode
 * :80-130): banner decides pattern / symmetric, 1-based entries to 0-based
 * COO in file order, values by strtod.  Returns the COO count (with the n
 * appended diagonal slots in adjacency mode) and fills n, m, symmetric;
 * with rows == NULL only counts.  Returns -1 on a malformed file. */
long orc_mm_read(const char *path, int adjacency, int *n, int *m, int *symmetric, int *rows,
                 int *cols, double *vals);
/* COO -> CSR (io_utils.c:141-235): rows counted, entries (and mirrors when
 * symmetric) filled in COO order, each row then stably merge-sorted by
 * column.  Returns the CSR entry count; with ja == NULL only counts (ia is
 * still filled). */
long orc_coo_to_csr(int n, long nnz, const int *rows, const int *cols, const double *vals,
                    int symmetric, int *ia, int *ja, double *a);
/* A = L + D + U (io_utils.c:268-363); augmented = 1 adds the diagonal last
 * in L and first in U (cg_driver2.c:95-144).  d (may be NULL) gets the
 * row's last diagonal entry or 0.  Returns nnz(L) and sets *nnz_u; with lja
 * == NULL only counts. */
long orc_split(int n, const int *ia, const int *ja, const double *a, int augmented, int *lia,
               int *lja, double *la, int *uia, int *uja, double *ua, double *d, long *nnz_u);
/* In place: adjacency (with diagonal slots) -> Laplacian (io_utils.c:
 * 383-425): int degree accumulated in CSR order from -1, unweighted or
 * normalised.  deg (may be NULL) receives the degrees. */
void orc_adjacency_laplacian(int n, const int *ia, const int *ja, double *a, int weighted,
                             int *deg);

#ifdef __cplusplus
}
#endif
#endif
