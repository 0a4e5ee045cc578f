/* Harness for the sanitizer builds of the oracle (lap_oracle.c): reads a
 * Matrix Market file with the oracle's reader, converts and splits it, runs
 * PCG with every preconditioner kind (and the OpenMP variant when asked),
 * and prints one line per solve: kind, iterations, flag, final residual
 * norm (%a, exact bits).  b[i] = sin(i + 1).
 *
 *   oracle_harness <matrix.mtx> [omp]                                     */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/lap_oracle.h"

static void *xcalloc(size_t count, size_t size) {
  void *p = calloc(count ? count : 1, size);
  if (!p) {
    fprintf(stderr, "oracle_harness: out of memory\n");
    exit(3);
  }
  return p;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: oracle_harness <matrix.mtx> [omp]\n");
    return 2;
  }
  const int omp = argc > 2 && strcmp(argv[2], "omp") == 0;
  int n = 0, m = 0, sym = 0;
  const long cnt = orc_mm_read(argv[1], 0, &n, &m, &sym, NULL, NULL, NULL);
  if (cnt < 0 || n <= 0) return 4;
  int *r = xcalloc((size_t)cnt, sizeof(int)), *c = xcalloc((size_t)cnt, sizeof(int));
  double *v = xcalloc((size_t)cnt, sizeof(double));
  orc_mm_read(argv[1], 0, &n, &m, &sym, r, c, v);
  int *ia = xcalloc((size_t)n + 1, sizeof(int));
  const long nnz = orc_coo_to_csr(n, cnt, r, c, v, sym, ia, NULL, NULL);
  int *ja = xcalloc((size_t)nnz, sizeof(int));
  double *a = xcalloc((size_t)nnz, sizeof(double));
  orc_coo_to_csr(n, cnt, r, c, v, sym, ia, ja, a);
  free(r);
  free(c);
  free(v);

  /* strict L / U with d (CPU path), and the augmented L+D / D+U for IC0 */
  int *lia = xcalloc((size_t)n + 1, sizeof(int)), *uia = xcalloc((size_t)n + 1, sizeof(int));
  long nu = 0;
  const long nl = orc_split(n, ia, ja, a, 0, lia, NULL, NULL, uia, NULL, NULL, NULL, &nu);
  int *lja = xcalloc((size_t)nl, sizeof(int)), *uja = xcalloc((size_t)nu, sizeof(int));
  double *la = xcalloc((size_t)nl, sizeof(double)), *ua = xcalloc((size_t)nu, sizeof(double));
  double *d = xcalloc((size_t)n, sizeof(double)), *d_r = xcalloc((size_t)n, sizeof(double));
  orc_split(n, ia, ja, a, 0, lia, lja, la, uia, uja, ua, d, &nu);
  orc_reciprocal(n, d, d_r);
  int *fia = xcalloc((size_t)n + 1, sizeof(int)), *gia = xcalloc((size_t)n + 1, sizeof(int));
  long ng = 0;
  const long nf = orc_split(n, ia, ja, a, 1, fia, NULL, NULL, gia, NULL, NULL, NULL, &ng);
  int *fja = xcalloc((size_t)nf, sizeof(int)), *gja = xcalloc((size_t)ng, sizeof(int));
  double *fa = xcalloc((size_t)nf, sizeof(double)), *ga = xcalloc((size_t)ng, sizeof(double));
  orc_split(n, ia, ja, a, 1, fia, fja, fa, gia, gja, ga, NULL, &ng);
  orc_ichol_factor(n, gia, gja, ga, fia, fja, fa);

  double *b = xcalloc((size_t)n, sizeof(double)), *x = xcalloc((size_t)n, sizeof(double));
  double *aux[3];
  for (int i = 0; i < 3; ++i) aux[i] = xcalloc((size_t)n, sizeof(double));
  for (int i = 0; i < n; ++i) b[i] = sin((double)(i + 1));
  const int maxit = 5000;
  double *hist = xcalloc((size_t)maxit + 2, sizeof(double));
  const int kinds[] = {ORC_PREC_NONE, ORC_PREC_LINE_JACOBI, ORC_PREC_IT_JACOBI, ORC_PREC_GS_STD,
                       ORC_PREC_GS_IT, ORC_PREC_GS_IT2, ORC_PREC_ICHOL};
  for (size_t q = 0; q < sizeof kinds / sizeof kinds[0]; ++q) {
    if (omp && kinds[q] != ORC_PREC_LINE_JACOBI) continue;
    const int ic = kinds[q] == ORC_PREC_ICHOL;
    orc_prec p = {kinds[q], n, 1, 2,
                  ic ? fia : lia, ic ? fja : lja, ic ? fa : la,
                  ic ? gia : uia, ic ? gja : uja, ic ? ga : ua,
                  d, d_r, aux[0], aux[1], aux[2]};
    memset(x, 0, sizeof(double) * (size_t)n);
    int it = 0, flag = 0;
    if (omp)
      orc_omp_cg(n, ia, ja, a, x, b, 1e-8, &p, maxit, &it, &flag, hist, 4, 4);
    else
      orc_cg(n, ia, ja, a, x, b, 1e-8, &p, maxit, &it, &flag, hist, 1);
    printf("kind %d it %d flag %d res %a\n", kinds[q], it, flag, hist[it]);
  }
  free(hist);
  for (int i = 0; i < 3; ++i) free(aux[i]);
  void *all[] = {b, x, ia, ja, a, lia, lja, la, uia, uja, ua, d, d_r, fia, fja, fa, gia, gja, ga};
  for (size_t i = 0; i < sizeof all / sizeof all[0]; ++i) free(all[i]);
  return 0;
}
