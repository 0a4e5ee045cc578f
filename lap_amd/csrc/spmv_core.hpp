// spmv_core.hpp -- the per-row-block SpMV bodies shared by the plain SpMV and
// the fused PCG kernels.
#pragma once

#include "lap_core.hpp"

namespace lap {

// Nonzeros staged in LDS per row block (8 B each -> 16 KiB).
constexpr int kSpmvTile = 2048;

// Long row (more than TILE nonzeros, a row block of its own): strided
// partial sums and a fixed-order tree over the workgroup.
template <int BS, class VT, class XT, class Epi>
__device__ __forceinline__ void spmv_long_row(int r0, int e0, int e1, const int *__restrict__ ja,
                                              const VT *__restrict__ a,
                                              const XT *__restrict__ x, double alpha,
                                              double *red, Epi &epi) {
  double s = 0.0;
#pragma unroll 4
  for (int e = e0 + (int)threadIdx.x; e < e1; e += BS)
    s += (alpha * (double)a[e]) * (double)x[ja[e]];
  const double t = block_sum<BS>(s, red);
  if (threadIdx.x == 0) epi.store(r0, epi.init(r0) + t);
}

// Row sums from the LDS products, in CSR order (row-sequential, as the
// reference's simple_csr_matvec, simple_blas.c:24-44).
template <int BS, class Epi>
__device__ __forceinline__ void spmv_row_sums(int r0, int r1, int e0, const int *__restrict__ ia,
                                              const double *prod, Epi &epi) {
  for (int row = r0 + (int)threadIdx.x; row < r1; row += BS) {
    const int s0 = ia[row] - e0, s1 = ia[row + 1] - e0;
    epi.store(row, lds_seq_sum(epi.init(row), prod, s0, s1));
  }
}

// The body for row block b: the block's (ja, a) read coalesced, x gathered
// and the products formed into LDS, then one lane per row sums them in CSR
// order.  Block-uniform control flow.  Values (VT) and vectors (XT) may be
// FP32; products and row sums are FP64.
//  prod : __shared__ double[TILE], red : __shared__ double[BS/64].
template <int BS, int TILE, class Epi, class VT = double, class XT = double>
__device__ __forceinline__ void spmv_block_rows(int b, const int *__restrict__ rowblk,
                                                const int *__restrict__ ia,
                                                const int *__restrict__ ja,
                                                const VT *__restrict__ a,
                                                const XT *__restrict__ x, double alpha,
                                                double *prod, double *red, Epi &epi) {
  const int r0 = rowblk[b], r1 = rowblk[b + 1];
  const int e0 = ia[r0], e1 = ia[r1];
  const int cnt = e1 - e0;
  if (cnt > TILE) {
    spmv_long_row<BS>(r0, e0, e1, ja, a, x, alpha, red, epi);
    return;
  }
#pragma unroll 4
  for (int k = threadIdx.x; k < cnt; k += BS) {
    const int e = e0 + k;
    prod[k] = (alpha * (double)a[e]) * (double)x[ja[e]];
  }
  __syncthreads();
  spmv_row_sums<BS>(r0, r1, e0, ia, prod, epi);
  __syncthreads();
}

// The body over value-indexed CSR: the value of nonzero e is
// dict[vi[e]] (dict in LDS; VT double, or float for the mixed-precision
// copy).  Same products and row order as spmv_block_rows.
template <int BS, int TILE, class Epi, class VT = double, class XT = double>
__device__ __forceinline__ void spmv_block_rows_vi(int b, const int *__restrict__ rowblk,
                                                   const int *__restrict__ ia,
                                                   const int *__restrict__ ja,
                                                   const unsigned char *__restrict__ vi,
                                                   const VT *dict,
                                                   const XT *__restrict__ x, double alpha,
                                                   double *prod, double *red, Epi &epi) {
  const int r0 = rowblk[b], r1 = rowblk[b + 1];
  const int e0 = ia[r0], e1 = ia[r1];
  const int cnt = e1 - e0;
  if (cnt > TILE) {
    double s = 0.0;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += BS)
      s += (alpha * (double)dict[vi[e]]) * (double)x[ja[e]];
    const double t = block_sum<BS>(s, red);
    if (threadIdx.x == 0) epi.store(r0, epi.init(r0) + t);
    return;
  }
#pragma unroll 4
  for (int k = threadIdx.x; k < cnt; k += BS) {
    const int e = e0 + k;
    prod[k] = (alpha * (double)dict[vi[e]]) * (double)x[ja[e]];
  }
  __syncthreads();
  spmv_row_sums<BS>(r0, r1, e0, ia, prod, epi);
  __syncthreads();
}

// A persistent walk over the row blocks of `w`.
template <int BS, int TILE, class Epi>
__device__ __forceinline__ void spmv_walk(BlockWalk w, const int *__restrict__ rowblk,
                                          const int *__restrict__ ia, const int *__restrict__ ja,
                                          const double *__restrict__ a,
                                          const double *__restrict__ x, double alpha,
                                          double *prod, double *red, Epi &epi) {
  for (int s = w.begin; s < w.end; s += w.step) {
    const int b = w.block(s);
    if (b >= w.nblk) continue;
    spmv_block_rows<BS, TILE>(b, rowblk, ia, ja, a, x, alpha, prod, red, epi);
  }
}

// spmv_walk over x + x2: each gathered value is x[c] + (double)x2[c], the
// value that folding x2 into x (x[c] += x2[c] in FP64) would have stored, so
// the sums are those of spmv_walk after such a fold, bit for bit.
template <int BS, int TILE, class X2T, class Epi>
__device__ __forceinline__ void spmv_walk_sum(BlockWalk w, const int *__restrict__ rowblk,
                                              const int *__restrict__ ia,
                                              const int *__restrict__ ja,
                                              const double *__restrict__ a,
                                              const double *__restrict__ x,
                                              const X2T *__restrict__ x2, double alpha,
                                              double *prod, double *red, Epi &epi) {
  for (int s = w.begin; s < w.end; s += w.step) {
    const int b = w.block(s);
    if (b >= w.nblk) continue;
    const int r0 = rowblk[b], r1 = rowblk[b + 1];
    const int e0 = ia[r0], e1 = ia[r1];
    const int cnt = e1 - e0;
    if (cnt > TILE) {  // a long row: as spmv_long_row
      double t = 0.0;
#pragma unroll 4
      for (int e = e0 + (int)threadIdx.x; e < e1; e += BS) {
        const int c = ja[e];
        t += (alpha * a[e]) * (x[c] + (double)x2[c]);
      }
      const double tt = block_sum<BS>(t, red);
      if (threadIdx.x == 0) epi.store(r0, epi.init(r0) + tt);
      continue;
    }
#pragma unroll 4
    for (int k = threadIdx.x; k < cnt; k += BS) {
      const int c = ja[e0 + k];
      prod[k] = (alpha * a[e0 + k]) * (x[c] + (double)x2[c]);
    }
    __syncthreads();
    spmv_row_sums<BS>(r0, r1, e0, ia, prod, epi);
    __syncthreads();
  }
}

}  // namespace lap
