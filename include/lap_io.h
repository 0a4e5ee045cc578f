/* lap_io.h -- input path of the gfx950 backend (liblapgfx.so).
 *
 * Host side: a parallel Matrix Market reader (lap_mm_load / lap_mm_coo), a
 * right-hand-side reader, a Matrix Market writer and a binary CSR format.
 * Device side: COO -> CSR, A = L + D + U (strict or with the diagonal
 * added to both triangles) and the graph Laplacian of an adjacency matrix.
 *
 * Every function returns LAP_OK (0) or an error code; lap_last_error()
 * (lap_gfx950.h) gives the message.  Device entry points take device
 * pointers and leave the library stream synchronised on return.
 *
 * The second half of this header defines the reference's io_utils.h
 * functions (read_mm_file, read_adjacency_file, coo_to_csr, read_rhs, split,
 * create_L_and_split) over a struct with the layout of its mmatrix, so a
 * reference driver links liblapgfx.so in place of io_utils.c.
 */
#ifndef LAP_IO_H
#define LAP_IO_H

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------- Matrix Market */

typedef struct lap_mm lap_mm; /* parsed file (COO in file order), host memory */

typedef struct {
  int n, m;      /* rows, columns from the size line */
  int nnz;       /* COO entries: file entries, + n diagonal slots in adjacency mode */
  int nnz_csr;   /* CSR entries after mirroring (nnz_unpacked of the reference) */
  int symmetric; /* banner says "symmetric", or adjacency mode */
  int pattern;   /* banner says "pattern": every value is 1.0 */
} lap_mm_info;

/* Parses a coordinate Matrix Market file with the semantics of
 * read_mm_file (io_utils.c:15-74): the banner decides "pattern" and
 * "symmetric", entries are 1-based "row col [value]" lines, values are
 * converted as sscanf("%lf") does (strtod), COO is 0-based in file order.
 * adjacency != 0 follows read_adjacency_file (io_utils.c:80-130): the file
 * is symmetric whatever its banner says and n entries (j, j, 1.0) are
 * appended.  Unlike the reference, comment and blank lines after the size
 * line are skipped, and a malformed line, an index out of range or an entry
 * count that differs from the size line is an error (LAP_ERR_STRUCT) rather
 * than undefined behaviour. */
int lap_mm_load(const char *path, int adjacency, lap_mm **out, lap_mm_info *info);
/* Copies the COO arrays (info.nnz each) to host (to_device = 0) or device
 * memory. */
int lap_mm_coo(const lap_mm *h, int *rows, int *cols, double *vals, int to_device);
void lap_mm_free(lap_mm *h);

/* Reads the first column of a Matrix Market "array" file into b[0..n)
 * (read_rhs, io_utils.c:239-264).  The size line's row count must equal n
 * (the reference leaves b[N..n) unset when it is smaller and writes past b
 * when it is larger). */
int lap_rhs_load(const char *path, int n, double *b, int to_device);

/* Writes the lower triangle (column <= row) of a host CSR matrix as
 * "coordinate real symmetric", values in %.17g (exact round trip). */
int lap_mm_write(const char *path, int n, const int *ia, const int *ja, const double *a);

/* ------------------------------------------------------------ binary CSR
 * File: 8-byte magic "LAPCSR\0\1", int64 n, int64 nnz, int64 0, then
 * ia int32[n+1], ja int32[nnz], a float64[nnz], little endian.  No
 * reference counterpart (the reference reads text only). */
int lap_csr_save(const char *path, int n, const int *ia, const int *ja, const double *a);
int lap_csr_info(const char *path, int *n, int *nnz);
/* Reads the arrays into host (to_device = 0) or device memory; the device
 * load streams the file through pinned buffers. */
int lap_csr_load(const char *path, int *ia, int *ja, double *a, int to_device);

/* ------------------------------------------------------------ on the device */

/* COO -> CSR (coo_to_csr, io_utils.c:141-235).  rows/cols/vals are nnz
 * device entries, 0-based, rows < n and cols < m (m == n when symmetric).
 * symmetric != 0 mirrors every off-diagonal entry.  Each CSR row holds its
 * entries sorted by column; entries with equal (row, column) keep the order
 * of the reference's fill (file order, a mirrored entry right after the
 * entries before it), which is the order of a stable sort of that fill
 * (glibc's qsort is a merge sort).  cap is the length of ja and a;
 * *nnz_csr receives the number of CSR entries, and LAP_ERR_ARG is returned
 * without writing ja/a when cap is too small (call with cap = 0 to query). */
int lap_coo_to_csr(int n, int m, int nnz, const int *rows, const int *cols, const double *vals,
                   int symmetric, int cap, int *ia, int *ja, double *a, int *nnz_csr);

/* A = L + D + U.  augmented = 0: strict triangles (split,
 * io_utils.c:268-363).  augmented = 1: L+D with the diagonal last in each
 * row and D+U with the diagonal first (the GS_std / IC0 set-up of
 * cg_driver2.c:95-196; a row without a stored diagonal gets 0.0).  d[i] is
 * the row's diagonal entry (the last one when stored twice, 0.0 when none);
 * d may be null.  cap_l / cap_u are the lengths of lja/la and uja/ua;
 * *nnz_l / *nnz_u receive the sizes (query with cap 0: only the counts are
 * written). */
int lap_csr_split(int n, const int *ia, const int *ja, const double *a, int augmented,
                  int cap_l, int *lia, int *lja, double *la, int cap_u, int *uia, int *uja,
                  double *ua, double *d, int *nnz_l, int *nnz_u);

/* In place: the adjacency matrix (with the diagonal slots added by the
 * adjacency reader) becomes the graph Laplacian, as create_L_and_split does
 * to A (io_utils.c:367-439).  deg[i] = sum of the row's stored values - 1,
 * accumulated in CSR order into an int (each partial sum truncated toward
 * zero, as the reference's int accumulator does).  weighted = 0: diagonal
 * entries become deg[i], off-diagonals -a.  weighted = 1: diagonal 1.0,
 * off-diagonal (i, j) becomes ((-a) * (1/sqrt(deg[i]))) * (1/sqrt(deg[j])).
 * deg (int[n], device) may be null. */
int lap_adjacency_laplacian(int n, const int *ia, const int *ja, double *a, int weighted,
                            int *deg);

/* -------------------------------------------- io_utils.h drop-in (host)
 * Same layout as the reference's mmatrix (io_utils.h:27-40); the names
 * below are the reference's, with its semantics (arrays allocated with
 * calloc, freed by the caller with free).  The parser is lap_mm_load's;
 * coo_to_csr, split and create_L_and_split run multi-threaded on the host
 * and produce the reference's arrays bit for bit.  On an unreadable or
 * malformed file read_mm_file / read_adjacency_file print the error to
 * stderr and leave n = nnz = 0 (the reference crashes or reads garbage).
 * Include the reference's io_utils.h or this block, not both. */
typedef struct {
  int *coo_rows;
  int *coo_cols;
  double *coo_vals;
  int *csr_ia;
  int *csr_ja;
  double *csr_vals;
  int n;
  int m;
  int nnz;
  int nnz_unpacked;
  int symmetric;
} lap_mmatrix;

void read_mm_file(const char *matrixFileName, lap_mmatrix *A);
void read_adjacency_file(const char *matrixFileName, lap_mmatrix *A);
void coo_to_csr(lap_mmatrix *A);
void read_rhs(const char *rhsFileName, double *rhs);
void split(lap_mmatrix *A, lap_mmatrix *L, lap_mmatrix *U, lap_mmatrix *D);
void create_L_and_split(lap_mmatrix *A, lap_mmatrix *L, lap_mmatrix *U, lap_mmatrix *D,
                        int weighted);

#ifdef __cplusplus
}
#endif

#endif /* LAP_IO_H */
