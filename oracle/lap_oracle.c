/*
 * lap_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A CPU restatement of the reference LAP solver's serial path, written from
 * the reference's behaviour (no source copied), used as the checker:
 *   - tests/ compare the HIP library against it on seeded inputs;
 *   - __graft_entry__.smoke() checks one small solve against it;
 *   - bench.py's cpu_baseline leg times orc_omp_cg when the compiled
 *     reference (oracle/_ref) is not present.
 * The product path (lap_amd/) never loads it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  Contraction
 * is off so a*b+c rounds twice, as in the reference's -std=c99 x86-64 build.
 *
 * Pinning (how each function is checked; tests/test_oracle.py):
 *   orc_dot/scal/axpy/vec_vec/reciprocal/sqrt  bitwise vs oracle/_ref
 *       (the compiled reference simple_blas.c) on random vectors, and
 *       closed forms (sum of 1..n, 1/x special cases).
 *   orc_csr_matvec    KAT test_spmv_omp.c:17-37 -> [-26.5,-26.5,-50.5,-32.5,-88.5];
 *       bitwise vs oracle/_ref; A*1 = 0 for an unshifted Laplacian; grid
 *       Laplacian cosine eigenvectors; dense numpy product.
 *   orc_lower_solve / orc_upper_solve   KAT test_tri_solve_omp.c:63-85;
 *       bitwise vs oracle/_ref; scipy solve_triangular.
 *   orc_ichol_factor  bitwise vs oracle/_ref initialize_ichol; invariant
 *       (L L^T)_ij = A_ij on pattern(A); exact Cholesky on a path graph.
 *   orc_ichol_apply   bitwise vs oracle/_ref simple_ichol.
 *   orc_prec_apply    bitwise vs oracle/_ref prec_function for every kind;
 *       dense numpy restatements; SGS closed form for GS_std with K=1.
 *   orc_cg            iteration count and residual history bitwise vs the
 *       reference's own cg() (oracle/_ref) and vs the committed golden
 *       fixtures produced by the reference binaries; 1-iteration convergence
 *       for an eigenvector RHS.
 *   orc_omp_*         same arithmetic as the serial path except the dot
 *       product's summation order; checked to agree within 1e-12.
 *   orc_graph_values  no reference counterpart: row sums L*1 = shift*1,
 *       symmetry, and a dense numpy assembly of sum_e w_e (u_s-u_d)(u_s-u_d)^T.
 *   orc_edge_grad     no reference counterpart: the gradient it gives with
 *       dense FP64 solves matches central finite differences of the loss
 *       (numpy, tests/test_oracle.py).
 */
#include "lap_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- BLAS-1 */

double orc_dot(int n, const double *v, const double *w) {
  double acc = 0.0;
  for (int i = 0; i < n; ++i) acc += v[i] * w[i];
  return acc;
}

void orc_scal(int n, double alpha, double *v) {
  for (int i = 0; i < n; ++i) v[i] *= alpha;
}

void orc_axpy(int n, double alpha, const double *x, double *y) {
  for (int i = 0; i < n; ++i) y[i] += alpha * x[i];
}

/* ---------------------------------------------------------------- sparse */

void orc_csr_matvec(int n, const int *ia, const int *ja, const double *a,
                    const double *x, double *y, double alpha, double beta) {
  for (int row = 0; row < n; ++row) {
    y[row] *= beta;
    for (int e = ia[row]; e < ia[row + 1]; ++e) y[row] += alpha * a[e] * x[ja[e]];
  }
}

/* One row of a triangular solve: subtract the off-diagonal terms in CSR
 * order, then divide by the diagonal (the in-row entry if present). */
static double tri_row(int row, const int *ia, const int *ja, const double *a,
                      const double *diag, const double *x, const double *y) {
  double acc = x[row];
  double piv = diag ? diag[row] : 0.0;
  for (int e = ia[row]; e < ia[row + 1]; ++e) {
    int c = ja[e];
    if (c == row) {
      piv = a[e];
      continue;
    }
    acc -= a[e] * y[c];
  }
  return acc / piv;
}

void orc_lower_solve(int n, const int *ia, const int *ja, const double *a,
                     const double *diag, const double *x, double *y) {
  for (int row = 0; row < n; ++row) y[row] = tri_row(row, ia, ja, a, diag, x, y);
}

void orc_upper_solve(int n, const int *ia, const int *ja, const double *a,
                     const double *diag, const double *x, double *y) {
  for (int row = n - 1; row >= 0; --row) y[row] = tri_row(row, ia, ja, a, diag, x, y);
}

void orc_vec_vec(int n, const double *x, const double *y, double *res) {
  for (int i = 0; i < n; ++i) res[i] = x[i] * y[i];
}

void orc_reciprocal(int n, const double *v, double *res) {
  for (int i = 0; i < n; ++i) res[i] = (v[i] != 0.0) ? 1.0 / v[i] : 0.0;
}

void orc_sqrt(int n, const double *v, double *res) {
  for (int i = 0; i < n; ++i) res[i] = (v[i] >= 0.0) ? sqrt(v[i]) : 0.0;
}

/* ------------------------------------------------------------------- IC0 */

void orc_ichol_factor(int n, const int *uia, const int *uja, double *ua,
                      const int *lia, const int *lja, double *la) {
  (void)lja;
  for (int i = 0; i < n; ++i) {
    const int head = uia[i], tail = uia[i + 1];
    ua[head] = sqrt(ua[head]);
    for (int e = head + 1; e < tail; ++e) ua[e] = ua[e] / ua[head];
    /* rank-1 update of the trailing rows touched by row i, restricted to
     * the sparsity pattern: U(r,c) -= U(i,r) U(i,c) for r <= c */
    for (int m = head + 1; m < tail; ++m) {
      const int r = uja[m];
      const double uir = ua[m];
      int k = uia[r], kend = uia[r + 1], l = m;
      while (k < kend && l < tail) {
        const int ck = uja[k], cl = uja[l];
        if (ck == cl) {
          ua[k] -= uir * ua[l];
          ++k;
          ++l;
        } else if (ck < cl) {
          ++k;
        } else {
          ++l;
        }
      }
    }
  }
  /* L = U^T: walking U by rows fills each L row in ascending column order */
  if (n <= 0) return;
  int *fill = (int *)calloc((size_t)n, sizeof(int));
  for (int i = 0; i < n; ++i)
    for (int e = uia[i]; e < uia[i + 1]; ++e) {
      const int r = uja[e];
      la[lia[r] + fill[r]++] = ua[e];
    }
  free(fill);
}

void orc_ichol_apply(int n, const int *lia, const int *lja, const double *la,
                     const int *uia, const int *uja, const double *ua,
                     const double *x, double *y, double *aux) {
  orc_lower_solve(n, lia, lja, la, NULL, x, aux);
  orc_upper_solve(n, uia, uja, ua, NULL, aux, y);
}

/* -------------------------------------------------------- preconditioners */

static void vcopy(int n, const double *s, double *d) { memcpy(d, s, sizeof(double) * (size_t)n); }
static void vzero(int n, double *v) { memset(v, 0, sizeof(double) * (size_t)n); }

/* K sweeps of y <- D^{-1}(x - (L+U) y) from y = 0 (it_jacobi.c:4-32) */
static void prec_it_jacobi(const orc_prec *p, const double *x, double *y) {
  const int n = p->n;
  vzero(n, y);
  for (int s = 0; s < p->k; ++s) {
    orc_csr_matvec(n, p->lia, p->lja, p->la, y, p->aux1, 1.0, 0.0);
    orc_csr_matvec(n, p->uia, p->uja, p->ua, y, p->aux2, 1.0, 0.0);
    orc_axpy(n, 1.0, p->aux1, p->aux2);
    vcopy(n, x, y);
    orc_axpy(n, -1.0, p->aux2, y);
    orc_vec_vec(n, y, p->d_r, y);
  }
}

/* K forward then K backward Gauss-Seidel corrections (GS.c:5-42) */
static void prec_gs_std(const int *ia, const int *ja, const double *a,
                        const orc_prec *p, const double *x, double *y) {
  const int n = p->n;
  vzero(n, y);
  for (int s = 0; s < p->k; ++s) {
    vcopy(n, x, p->aux2);
    orc_csr_matvec(n, ia, ja, a, y, p->aux2, -1.0, 1.0);
    orc_lower_solve(n, p->lia, p->lja, p->la, p->d, p->aux2, p->aux1);
    orc_axpy(n, 1.0, p->aux1, y);
  }
  for (int s = 0; s < p->k; ++s) {
    vcopy(n, x, p->aux2);
    orc_csr_matvec(n, ia, ja, a, y, p->aux2, -1.0, 1.0);
    orc_upper_solve(n, p->uia, p->uja, p->ua, p->d, p->aux2, p->aux1);
    orc_axpy(n, 1.0, p->aux1, y);
  }
}

/* M outer corrections, each approximating the two triangular solves by K
 * Jacobi-style inner sweeps (GS.c:45-92) */
static void prec_gs_it(const int *ia, const int *ja, const double *a,
                       const orc_prec *p, const double *x, double *y) {
  const int n = p->n;
  vzero(n, y);
  for (int o = 0; o < p->m; ++o) {
    vcopy(n, x, p->aux2);
    orc_csr_matvec(n, ia, ja, a, y, p->aux2, -1.0, 1.0);
    orc_vec_vec(n, p->aux2, p->d_r, p->aux1);
    for (int s = 0; s < p->k; ++s) {
      vcopy(n, p->aux2, p->aux3);
      orc_csr_matvec(n, p->lia, p->lja, p->la, p->aux1, p->aux3, -1.0, 1.0);
      orc_vec_vec(n, p->aux3, p->d_r, p->aux1);
    }
    for (int s = 0; s < p->k; ++s) {
      vcopy(n, p->aux2, p->aux3);
      orc_csr_matvec(n, p->uia, p->uja, p->ua, p->aux1, p->aux3, -1.0, 1.0);
      orc_vec_vec(n, p->aux3, p->d_r, p->aux1);
    }
    orc_axpy(n, 1.0, p->aux1, y);
  }
}

/* Second iterative GS variant (GS.c:95-159) */
static void prec_gs_it2(const orc_prec *p, const double *x, double *y) {
  const int n = p->n;
  orc_vec_vec(n, x, p->d_r, p->aux1);
  for (int o = 0; o < p->m; ++o) {
    orc_csr_matvec(n, p->lia, p->lja, p->la, p->aux1, p->aux2, 1.0, 0.0);
    orc_csr_matvec(n, p->uia, p->uja, p->ua, p->aux1, p->aux3, 1.0, 0.0);
    orc_axpy(n, 1.0, p->aux3, p->aux2);
    vcopy(n, x, p->aux3);
    orc_axpy(n, -1.0, p->aux2, p->aux3);
    orc_vec_vec(n, p->aux3, p->d_r, p->aux2);

    vcopy(n, x, p->aux3);
    orc_csr_matvec(n, p->lia, p->lja, p->la, p->aux2, p->aux3, -1.0, 1.0);
    for (int s = 0; s < p->k; ++s) {
      vcopy(n, p->aux3, p->aux1);
      orc_csr_matvec(n, p->uia, p->uja, p->ua, p->aux2, p->aux1, -1.0, 1.0);
      orc_vec_vec(n, p->d_r, p->aux1, p->aux2);
    }
    vcopy(n, x, p->aux3);
    orc_csr_matvec(n, p->uia, p->uja, p->ua, p->aux2, p->aux3, -1.0, 1.0);
    for (int s = 0; s < p->k; ++s) {
      vcopy(n, p->aux3, p->aux1);
      orc_csr_matvec(n, p->lia, p->lja, p->la, p->aux2, p->aux1, -1.0, 1.0);
      orc_vec_vec(n, p->d_r, p->aux1, p->aux2);
    }
    vcopy(n, p->aux2, p->aux1);
  }
  vcopy(n, p->aux2, y);
}

/* string dispatch of prec.c:6-45, by enum */
void orc_prec_apply(const int *ia, const int *ja, const double *a,
                    const orc_prec *p, const double *x, double *y) {
  switch (p->kind) {
    case ORC_PREC_LINE_JACOBI: orc_vec_vec(p->n, p->d_r, x, y); break;
    case ORC_PREC_IT_JACOBI: prec_it_jacobi(p, x, y); break;
    case ORC_PREC_GS_STD: prec_gs_std(ia, ja, a, p, x, y); break;
    case ORC_PREC_GS_IT: prec_gs_it(ia, ja, a, p, x, y); break;
    case ORC_PREC_GS_IT2: prec_gs_it2(p, x, y); break;
    case ORC_PREC_ICHOL:
      orc_ichol_apply(p->n, p->lia, p->lja, p->la, p->uia, p->uja, p->ua, x, y, p->aux1);
      break;
    default: vcopy(p->n, x, y); break;
  }
}

/* -------------------------------------------------------------------- PCG */

typedef double (*dot_fn)(int, const double *, const double *, int);
typedef void (*mv_fn)(int, const int *, const int *, const double *, const double *,
                      double *, double, double, int);

static double dot_seq(int n, const double *v, const double *w, int t) {
  (void)t;
  return orc_dot(n, v, w);
}
static void mv_seq(int n, const int *ia, const int *ja, const double *a, const double *x,
                   double *y, double al, double be, int t) {
  (void)t;
  orc_csr_matvec(n, ia, ja, a, x, y, al, be);
}

static double dot_omp(int n, const double *v, const double *w, int threads) {
  double acc = 0.0;
#pragma omp parallel for num_threads(threads) reduction(+ : acc) schedule(static)
  for (int i = 0; i < n; ++i) acc += v[i] * w[i];
  return acc;
}

void orc_omp_csr_matvec(int n, const int *ia, const int *ja, const double *a,
                        const double *x, double *y, double alpha, double beta,
                        int threads) {
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int row = 0; row < n; ++row) {
    double acc = y[row] * beta;
    for (int e = ia[row]; e < ia[row + 1]; ++e) acc += alpha * a[e] * x[ja[e]];
    y[row] = acc;
  }
}

static void cg_body(int n, const int *ia, const int *ja, const double *a, double *x,
                    const double *b, double tol, const orc_prec *p, int maxit, int *it,
                    int *flag, double *hist, int quiet, dot_fn dot, mv_fn mv,
                    int mv_threads, int dot_threads) {
  double *r = (double *)calloc((size_t)n, sizeof(double));
  double *w = (double *)calloc((size_t)n, sizeof(double));
  double *pp = (double *)calloc((size_t)n, sizeof(double));
  double *q = (double *)calloc((size_t)n, sizeof(double));
  double rho = 0.0, rho_prev = 0.0;
  int iter = 0;

  /* r = -(A x - b) */
  mv(n, ia, ja, a, x, r, 1.0, 0.0, mv_threads);
  orc_axpy(n, -1.0, b, r);
  orc_scal(n, -1.0, r);
  hist[0] = sqrt(dot(n, r, r, dot_threads));
  const double tolrel = tol * hist[0];
  if (!quiet) printf("CG: it %d, res norm %5.5e \n", 0, hist[0]);

  for (;;) {
    orc_prec_apply(ia, ja, a, p, r, w);
    rho = dot(n, r, w, dot_threads);
    if (iter == 0) {
      vcopy(n, w, pp);
    } else {
      const double beta = rho / rho_prev;
      orc_scal(n, beta, pp);
      orc_axpy(n, 1.0, w, pp);
    }
    mv(n, ia, ja, a, pp, q, 1.0, 0.0, mv_threads);
    const double alpha = rho / dot(n, pp, q, dot_threads);
    orc_axpy(n, alpha, pp, x);
    orc_axpy(n, -1.0 * alpha, q, r);
    ++iter;
    hist[iter] = sqrt(dot(n, r, r, dot_threads));
    if (!quiet) printf("CG: it %d, res norm %5.5e \n", iter, hist[iter]);
    if (hist[iter] < tolrel) {
      *flag = 0;
      *it = iter;
      if (!quiet) {
        mv(n, ia, ja, a, x, r, 1.0, 0.0, mv_threads);
        orc_axpy(n, -1.0, b, r);
        const double tr = sqrt(dot(n, r, r, dot_threads));
        printf("TRUE Norm of r %5.5e relative %16.16e\n", tr,
               tr / sqrt(dot(n, b, b, dot_threads)));
      }
      break;
    }
    if (iter > maxit) {
      *flag = 1;
      *it = iter;
      break;
    }
    rho_prev = rho;
  }
  free(r);
  free(w);
  free(pp);
  free(q);
}

void orc_cg(int n, const int *ia, const int *ja, const double *a, double *x,
            const double *b, double tol, const orc_prec *p, int maxit, int *it,
            int *flag, double *hist, int quiet) {
  cg_body(n, ia, ja, a, x, b, tol, p, maxit, it, flag, hist, quiet, dot_seq, mv_seq, 1, 1);
}

static void mv_omp(int n, const int *ia, const int *ja, const double *a, const double *x,
                   double *y, double al, double be, int t) {
  orc_omp_csr_matvec(n, ia, ja, a, x, y, al, be, t);
}

void orc_omp_cg(int n, const int *ia, const int *ja, const double *a, double *x,
                const double *b, double tol, const orc_prec *p, int maxit, int *it,
                int *flag, double *hist, int spmv_threads, int dot_threads) {
  cg_body(n, ia, ja, a, x, b, tol, p, maxit, it, flag, hist, 1, dot_omp, mv_omp,
          spmv_threads, dot_threads);
}

/* ------------------------------------------- weighted graph Laplacian */

void orc_graph_values(int n, int m, const int *ia, const int *ja, const int *epos,
                      const int *dpos, const double *w, double shift, double *a) {
  for (int e = 0; e < m; ++e) {
    a[epos[2 * e]] = -w[e];
    a[epos[2 * e + 1]] = -w[e];
  }
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int p = ia[i]; p < ia[i + 1]; ++p)
      if (ja[p] != i) s += -a[p];
    a[dpos[i]] = s + shift;
  }
}

void orc_edge_grad(int m, const int *src, const int *dst, const double *lam, const double *x,
                   double *g) {
  for (int e = 0; e < m; ++e) g[e] = -((lam[src[e]] - lam[dst[e]]) * (x[src[e]] - x[dst[e]]));
}

/* ------------------------------------------------------------ input path */

long orc_mm_read(const char *path, int adjacency, int *n, int *m, int *symmetric, int *rows,
                 int *cols, double *vals) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char line[512];
  if (!fgets(line, sizeof line, f)) {
    fclose(f);
    return -1;
  }
  const int pattern = strstr(line, "pattern") != NULL;
  const int sym = adjacency || strstr(line, "symmetric") != NULL;
  while (line[0] == '%')
    if (!fgets(line, sizeof line, f)) {
      fclose(f);
      return -1;
    }
  long nn, mm, nz;
  if (sscanf(line, "%ld %ld %ld", &nn, &mm, &nz) != 3) {
    fclose(f);
    return -1;
  }
  long k = 0;
  while (k < nz && fgets(line, sizeof line, f)) {
    char *p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '%' || *p == '\n' || *p == '\r' || *p == 0) continue;
    char *e;
    const long r = strtol(p, &e, 10);
    const long c = strtol(e, &e, 10);
    double v = 1.0;
    if (!pattern) v = strtod(e, &e);
    if (rows) {
      rows[k] = (int)(r - 1);
      cols[k] = (int)(c - 1);
      vals[k] = v;
    }
    ++k;
  }
  fclose(f);
  if (k != nz) return -1;
  if (adjacency) {
    for (long j = 0; j < nn; ++j, ++k)
      if (rows) {
        rows[k] = (int)j;
        cols[k] = (int)j;
        vals[k] = 1.0;
      }
  }
  *n = (int)nn;
  *m = (int)mm;
  *symmetric = sym;
  return k;
}

/* stable bottom-up merge sort of (key, val) by key */
static void merge_sort_row(int len, int *key, double *val, int *tk, double *tv) {
  for (int w = 1; w < len; w *= 2) {
    for (int lo = 0; lo < len; lo += 2 * w) {
      const int mid = lo + w < len ? lo + w : len, hi = lo + 2 * w < len ? lo + 2 * w : len;
      int i = lo, j = mid, o = lo;
      while (i < mid && j < hi) {
        if (key[j] < key[i]) {
          tk[o] = key[j];
          tv[o++] = val[j++];
        } else {
          tk[o] = key[i];
          tv[o++] = val[i++];
        }
      }
      while (i < mid) {
        tk[o] = key[i];
        tv[o++] = val[i++];
      }
      while (j < hi) {
        tk[o] = key[j];
        tv[o++] = val[j++];
      }
    }
    memcpy(key, tk, sizeof(int) * (size_t)len);
    memcpy(val, tv, sizeof(double) * (size_t)len);
  }
}

long orc_coo_to_csr(int n, long nnz, const int *rows, const int *cols, const double *vals,
                    int symmetric, int *ia, int *ja, double *a) {
  long *cnt = calloc((size_t)n + 1, sizeof(long));
  for (long i = 0; i < nnz; ++i) {
    cnt[rows[i]]++;
    if (symmetric && rows[i] != cols[i]) cnt[cols[i]]++;
  }
  ia[0] = 0;
  for (int i = 0; i < n; ++i) ia[i + 1] = ia[i] + (int)cnt[i];
  const long total = ia[n];
  if (ja) {
    for (int i = 0; i < n; ++i) cnt[i] = ia[i];
    for (long i = 0; i < nnz; ++i) {
      const long p = cnt[rows[i]]++;
      ja[p] = cols[i];
      a[p] = vals[i];
      if (symmetric && rows[i] != cols[i]) {
        const long q = cnt[cols[i]]++;
        ja[q] = rows[i];
        a[q] = vals[i];
      }
    }
    int maxlen = 1;
    for (int i = 0; i < n; ++i)
      if (ia[i + 1] - ia[i] > maxlen) maxlen = ia[i + 1] - ia[i];
    int *tk = malloc(sizeof(int) * (size_t)maxlen);
    double *tv = malloc(sizeof(double) * (size_t)maxlen);
    for (int i = 0; i < n; ++i) merge_sort_row(ia[i + 1] - ia[i], ja + ia[i], a + ia[i], tk, tv);
    free(tk);
    free(tv);
  }
  free(cnt);
  return total;
}

long orc_split(int n, const int *ia, const int *ja, const double *a, int augmented, int *lia,
               int *lja, double *la, int *uia, int *uja, double *ua, double *d, long *nnz_u) {
  long nl = 0, nu = 0;
  for (int i = 0; i < n; ++i) {
    lia[i] = (int)nl;
    uia[i] = (int)nu;
    double dv = 0.0;
    long pu = nu + augmented;
    for (int p = ia[i]; p < ia[i + 1]; ++p) {
      const int c = ja[p];
      if (c < i) {
        if (lja) {
          lja[nl] = c;
          la[nl] = a[p];
        }
        ++nl;
      } else if (c > i) {
        if (uja) {
          uja[pu] = c;
          ua[pu] = a[p];
        }
        ++pu;
      } else {
        dv = a[p];
      }
    }
    if (augmented) {
      if (lja) {
        lja[nl] = i;
        la[nl] = dv;
        uja[nu] = i;
        ua[nu] = dv;
      }
      ++nl;
    }
    nu = pu;
    if (d) d[i] = dv;
  }
  lia[n] = (int)nl;
  uia[n] = (int)nu;
  *nnz_u = nu;
  return nl;
}

void orc_adjacency_laplacian(int n, const int *ia, const int *ja, double *a, int weighted,
                             int *deg) {
  int *dd = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) {
    dd[i] = -1;
    for (int p = ia[i]; p < ia[i + 1]; ++p) dd[i] += a[p]; /* int += double truncates */
  }
  for (int i = 0; i < n; ++i) {
    const double ds = weighted ? 1.0 / sqrt(dd[i]) : 0.0;
    for (int p = ia[i]; p < ia[i + 1]; ++p) {
      if (ja[p] == i)
        a[p] = weighted ? 1.0 : (double)dd[i];
      else if (!weighted)
        a[p] = (-1.0) * a[p];
      else
        a[p] = (-1.0) * a[p] * ds * (1.0 / sqrt(dd[ja[p]]));
    }
  }
  if (deg) memcpy(deg, dd, sizeof(int) * (size_t)n);
  free(dd);
}
