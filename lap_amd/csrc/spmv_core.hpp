// spmv_core.hpp -- the per-row-block SpMV bodies shared by the plain SpMV and
// the fused PCG kernels.
#pragma once

#include "lap_core.hpp"

namespace lap {

// Nonzeros staged in LDS per row block (8 B each -> 16 KiB).
constexpr int kSpmvTile = 2048;

// Variants (selected per launch, results bitwise identical):
//  0  LDS stream: strided loop of (ja, a, x-gather) into LDS, then row sums.
//  1  register-staged: every lane issues all of its (ja, a) loads for the
//     row block at once (TILE/BS per lane), then all x gathers, then the LDS
//     products; the next row block's bounds are fetched while the current
//     one is summed.
//  2  as 1 with non-temporal (streaming) loads of ja and a, which are read
//     exactly once, so they do not evict reused x entries from L2.
//  3  software-pipelined: the next row block's (ja, a) loads are issued
//     right after the current block's x gathers, so they are in flight
//     while the current block is summed (barriers on gfx950 wait for LDS
//     only, not for outstanding global loads).
//  4  as 3 with non-temporal loads of ja and a.
//  5  as 0 with non-temporal loads of ja and a (read once, so they do not
//     push the reused x lines out of the XCD's L2).
//  6  as 0 with a shorter chain of dependent loads per block: the block's
//     nonzero range comes from the plan (no ia load before the ja/a loads),
//     and the rows' pointers are staged into LDS while the products are
//     formed (no ia load in the row-sum phase).
constexpr int kSpmvVariants = 7;

template <bool NT, class T>
__device__ __forceinline__ T ld_stream(const T *p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

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

// Variant 0 body for row block b.  Block-uniform control flow.  Values
// (VT) and vectors (XT) may be FP32; products and row sums are FP64.
//  prod : __shared__ double[TILE], red : __shared__ double[BS/64].
template <int BS, int TILE, class Epi, class VT = double, class XT = double, bool NT = false>
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
    prod[k] = (alpha * (double)ld_stream<NT>(a + e)) * (double)x[ld_stream<NT>(ja + e)];
  }
  __syncthreads();
  spmv_row_sums<BS>(r0, r1, e0, ia, prod, epi);
  __syncthreads();
}

// Variant 6 body: e0/e1 from the plan's offsets `ebl`, row pointers in LDS.
template <int BS, int TILE, class Epi>
__device__ __forceinline__ void spmv_block_rows_pf(int b, const int *__restrict__ rowblk,
                                                   const int *__restrict__ ebl,
                                                   const int *__restrict__ ia,
                                                   const int *__restrict__ ja,
                                                   const double *__restrict__ a,
                                                   const double *__restrict__ x, double alpha,
                                                   double *prod, double *red, Epi &epi) {
  __shared__ int rp[2 * BS + 1];  // row pointers relative to e0 (blocks hold <= 2*BS rows)
  const int r0 = rowblk[b], r1 = rowblk[b + 1];
  const int e0 = ebl[b], e1 = ebl[b + 1];
  const int cnt = e1 - e0;
  if (cnt > TILE) {
    spmv_long_row<BS>(r0, e0, e1, ja, a, x, alpha, red, epi);
    return;
  }
  for (int i = threadIdx.x; i <= r1 - r0; i += BS) rp[i] = ia[r0 + i] - e0;
#pragma unroll 4
  for (int k = threadIdx.x; k < cnt; k += BS) {
    const int e = e0 + k;
    prod[k] = (alpha * a[e]) * x[ja[e]];
  }
  __syncthreads();
  for (int row = r0 + (int)threadIdx.x; row < r1; row += BS) {
    const int s0 = rp[row - r0], s1 = rp[row - r0 + 1];
    epi.store(row, lds_seq_sum(epi.init(row), prod, s0, s1));
  }
  __syncthreads();
}

// Variant 0 body over value-indexed CSR: the value of nonzero e is
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

// Variant 0 body with x given by a functor xf(j) (a value computed from
// several vectors at the gathered index).  Same products and row order as
// spmv_block_rows with alpha = 1.
template <int BS, int TILE, class Epi, class XF>
__device__ __forceinline__ void spmv_block_rows_fx(int b, const int *__restrict__ rowblk,
                                                   const int *__restrict__ ia,
                                                   const int *__restrict__ ja,
                                                   const double *__restrict__ a, const XF &xf,
                                                   double *prod, double *red, Epi &epi) {
  const int r0 = rowblk[b], r1 = rowblk[b + 1];
  const int e0 = ia[r0], e1 = ia[r1];
  const int cnt = e1 - e0;
  if (cnt > TILE) {
    double s = 0.0;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += BS) s += (1.0 * a[e]) * xf(ja[e]);
    const double t = block_sum<BS>(s, red);
    if (threadIdx.x == 0) epi.store(r0, epi.init(r0) + t);
    return;
  }
#pragma unroll 4
  for (int k = threadIdx.x; k < cnt; k += BS) {
    const int e = e0 + k;
    prod[k] = (1.0 * a[e]) * xf(ja[e]);
  }
  __syncthreads();
  spmv_row_sums<BS>(r0, r1, e0, ia, prod, epi);
  __syncthreads();
}

// Variants 1/2: a persistent walk over the row blocks [w.begin, w.end) with
// stride w.step, register-staged loads and prefetched block bounds.
template <int BS, int TILE, bool NT, class Epi>
__device__ __forceinline__ void spmv_walk_staged(BlockWalk w, const int *__restrict__ rowblk,
                                                 const int *__restrict__ ia,
                                                 const int *__restrict__ ja,
                                                 const double *__restrict__ a,
                                                 const double *__restrict__ x, double alpha,
                                                 double *prod, double *red, Epi &epi) {
  constexpr int E = TILE / BS;
  int b = w.begin;
  if (b >= w.end) return;
  int r0 = rowblk[b], r1 = rowblk[b + 1];
  int e0 = ia[r0], e1 = ia[r1];
  while (true) {
    const int cnt = e1 - e0;
    const int nb = b + w.step;
    const bool more = nb < w.end;
    int nr0 = 0, nr1 = 0;
    if (more) {  // next block's bounds, fetched while this block runs
      nr0 = rowblk[nb];
      nr1 = rowblk[nb + 1];
    }
    if (cnt > TILE) {
      spmv_long_row<BS>(r0, e0, e1, ja, a, x, alpha, red, epi);
    } else {
      int jv[E];
      double av[E];
#pragma unroll
      for (int u = 0; u < E; ++u) {
        const int k = (int)threadIdx.x + u * BS;
        if (k < cnt) {
          jv[u] = ld_stream<NT>(ja + e0 + k);
          av[u] = ld_stream<NT>(a + e0 + k);
        }
      }
#pragma unroll
      for (int u = 0; u < E; ++u) {
        const int k = (int)threadIdx.x + u * BS;
        if (k < cnt) prod[k] = (alpha * av[u]) * x[jv[u]];
      }
      __syncthreads();
      spmv_row_sums<BS>(r0, r1, e0, ia, prod, epi);
      __syncthreads();
    }
    if (!more) break;
    b = nb;
    r0 = nr0;
    r1 = nr1;
    e0 = ia[r0];
    e1 = ia[r1];
  }
}

// Variants 3/4: software-pipelined walk (one register set, refilled with the
// next block's ja/a as soon as the current block's gathers consumed it).
template <int BS, int TILE, bool NT, class Epi>
__device__ __forceinline__ void spmv_walk_pipelined(BlockWalk w, const int *__restrict__ rowblk,
                                                    const int *__restrict__ ia,
                                                    const int *__restrict__ ja,
                                                    const double *__restrict__ a,
                                                    const double *__restrict__ x, double alpha,
                                                    double *prod, double *red, Epi &epi) {
  constexpr int E = TILE / BS;
  int b = w.begin;
  if (b >= w.end) return;
  int jv[E];
  double av[E];
  int r0 = rowblk[b], r1 = rowblk[b + 1];
  int e0 = ia[r0], e1 = ia[r1];
  bool loaded = false;
  while (true) {
    const int cnt = e1 - e0;
    if (!loaded && cnt <= TILE) {
#pragma unroll
      for (int u = 0; u < E; ++u) {
        const int k = (int)threadIdx.x + u * BS;
        if (k < cnt) {
          jv[u] = ld_stream<NT>(ja + e0 + k);
          av[u] = ld_stream<NT>(a + e0 + k);
        }
      }
    }
    loaded = false;
    const int nb = b + w.step;
    const bool more = nb < w.end;
    int nr0 = 0, nr1 = 0, ne0 = 0, ne1 = 0;
    if (more) {
      nr0 = rowblk[nb];
      nr1 = rowblk[nb + 1];
      ne0 = ia[nr0];
      ne1 = ia[nr1];
    }
    if (cnt > TILE) {
      spmv_long_row<BS>(r0, e0, e1, ja, a, x, alpha, red, epi);
    } else {
#pragma unroll
      for (int u = 0; u < E; ++u) {
        const int k = (int)threadIdx.x + u * BS;
        if (k < cnt) prod[k] = (alpha * av[u]) * x[jv[u]];
      }
      const int ncnt = ne1 - ne0;
      if (more && ncnt <= TILE) {
#pragma unroll
        for (int u = 0; u < E; ++u) {
          const int k = (int)threadIdx.x + u * BS;
          if (k < ncnt) {
            jv[u] = ld_stream<NT>(ja + ne0 + k);
            av[u] = ld_stream<NT>(a + ne0 + k);
          }
        }
        loaded = true;
      }
      __syncthreads();
      spmv_row_sums<BS>(r0, r1, e0, ia, prod, epi);
      __syncthreads();
    }
    if (!more) break;
    b = nb;
    r0 = nr0;
    r1 = nr1;
    e0 = ne0;
    e1 = ne1;
  }
}

// Dispatch over the variants for a persistent walk.
template <int BS, int TILE, int V, class Epi>
__device__ __forceinline__ void spmv_walk(BlockWalk w, const int *__restrict__ rowblk,
                                          const int *__restrict__ ia, const int *__restrict__ ja,
                                          const double *__restrict__ a,
                                          const double *__restrict__ x, double alpha,
                                          double *prod, double *red, Epi &epi) {
  if (V == 0 || V == 5 || (w.g && V != 6)) {  // the staged walks need a plain stride
    for (int s = w.begin; s < w.end; s += w.step) {
      const int b = w.block(s);
      if (b >= w.nblk) continue;
      spmv_block_rows<BS, TILE, Epi, double, double, V == 5>(b, rowblk, ia, ja, a, x, alpha, prod,
                                                             red, epi);
    }
  } else if constexpr (V == 6) {
    const int *ebl = rowblk + w.nblk + 1;
    for (int s = w.begin; s < w.end; s += w.step) {
      const int b = w.block(s);
      if (b >= w.nblk) continue;
      spmv_block_rows_pf<BS, TILE>(b, rowblk, ebl, ia, ja, a, x, alpha, prod, red, epi);
    }
  } else if constexpr (V <= 2) {
    spmv_walk_staged<BS, TILE, V == 2>(w, rowblk, ia, ja, a, x, alpha, prod, red, epi);
  } else {
    spmv_walk_pipelined<BS, TILE, V == 4>(w, rowblk, ia, ja, a, x, alpha, prod, red, epi);
  }
}

}  // namespace lap
