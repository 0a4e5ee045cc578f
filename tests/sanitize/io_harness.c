/* Host-only harness for the sanitizer builds of the input path: reads a
 * Matrix Market file with the library's io_utils.h drop-in (parallel parser,
 * atomic COO -> CSR, split or create_L_and_split) and with lap_mm_load /
 * lap_mm_coo, and writes every resulting array to a binary file that the test
 * compares with the regular build's.  No GPU call is made.
 *
 *   io_harness mm|adj <matrix.mtx> <out.bin> [weighted (adj only)]        */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lap_io.h"

static void put(FILE *f, const void *p, size_t bytes) {
  if (bytes && fwrite(p, 1, bytes, f) != bytes) {
    fprintf(stderr, "io_harness: short write\n");
    exit(3);
  }
}

static void put_csr(FILE *f, const lap_mmatrix *M, int n, int nnz) {
  put(f, M->csr_ia, sizeof(int) * (size_t)(n + 1));
  put(f, M->csr_ja, sizeof(int) * (size_t)nnz);
  put(f, M->csr_vals, sizeof(double) * (size_t)nnz);
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: io_harness mm|adj <matrix.mtx> <out.bin> [weighted]\n");
    return 2;
  }
  const int adj = strcmp(argv[1], "adj") == 0;
  const int weighted = argc > 4 ? atoi(argv[4]) : -1;
  lap_mmatrix A, L, U, D;
  memset(&A, 0, sizeof A);
  memset(&L, 0, sizeof L);
  memset(&U, 0, sizeof U);
  memset(&D, 0, sizeof D);
  if (adj)
    read_adjacency_file(argv[2], &A);
  else
    read_mm_file(argv[2], &A);
  if (A.n <= 0) return 4;
  const int nnz_coo = A.nnz;
  coo_to_csr(&A);
  if (adj && weighted >= 0)
    create_L_and_split(&A, &L, &U, &D, weighted);
  else
    split(&A, &L, &U, &D);

  /* the handle API on the same file: COO in file order, host copies */
  lap_mm *h = NULL;
  lap_mm_info info;
  if (lap_mm_load(argv[2], adj, &h, &info) != 0) return 5;
  int *r = malloc(sizeof(int) * (size_t)(info.nnz + 1));
  int *c = malloc(sizeof(int) * (size_t)(info.nnz + 1));
  double *v = malloc(sizeof(double) * (size_t)(info.nnz + 1));
  if (lap_mm_coo(h, r, c, v, 0) != 0) return 6;
  lap_mm_free(h);

  FILE *f = fopen(argv[3], "wb");
  if (!f) return 7;
  const int head[10] = {A.n, A.m, nnz_coo, A.nnz_unpacked, A.symmetric, L.nnz, U.nnz, D.nnz,
                        info.nnz, info.nnz_csr};
  put(f, head, sizeof head);
  put_csr(f, &A, A.n, A.nnz_unpacked);
  put_csr(f, &L, A.n, L.nnz);
  put_csr(f, &U, A.n, U.nnz);
  put(f, D.csr_vals, sizeof(double) * (size_t)A.n);
  put(f, r, sizeof(int) * (size_t)info.nnz);
  put(f, c, sizeof(int) * (size_t)info.nnz);
  put(f, v, sizeof(double) * (size_t)info.nnz);
  fclose(f);
  free(r);
  free(c);
  free(v);
  lap_mmatrix *all[4] = {&A, &L, &U, &D};
  for (int i = 0; i < 4; ++i) {
    free(all[i]->csr_ia);
    free(all[i]->csr_ja);
    free(all[i]->csr_vals);
    free(all[i]->coo_rows);
    free(all[i]->coo_cols);
    free(all[i]->coo_vals);
  }
  return 0;
}
