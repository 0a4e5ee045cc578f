// mixed.hip -- mixed-precision PCG (lap_pcg_mixed): an FP32 inner PCG inside
// FP64 iterative refinement.
//
// Refinement scheme ("reliable updates"): the inner PCG runs on FP32 vectors
// (r32, p32, q32, a correction x32) with the FP32 copy of A and of the
// preconditioner.  When its recursive residual has dropped by `delta`
// relative to the last refinement point (or looks converged, or maxit is
// reached), the FP32 correction is folded into the FP64 solution
// (x += x32, x32 = 0), the residual is recomputed in FP64 (r = b - A x with
// the caller's FP64 matrix) and replaces r32.  The Krylov process is not
// restarted (p and rho carry over), so the iteration count stays close to
// the FP64 solver's, whereas a restarted inner solve per refinement step
// discards the Krylov space every time.  Convergence (||b - A x|| <
// tol*||r0||) is only ever declared on an FP64 residual.
//
// FP64 is used for: the caller's x and b, the refinement residual r = b - A x
// (FP64 CSR SpMV), the diagonal of A (a shifted graph Laplacian's diagonal
// degree + 1e-8 does not survive FP32 rounding, which would make the FP32
// matrix singular), every row sum and every dot product.  FP32 holds: the
// off-diagonal values of A (own CSR without the diagonal), all inner
// vectors, 1/diag for Jacobi and the IC0 factor.
//
// Row-sum-preserving rounding: the inner operator's FP64 diagonal is A's
// diagonal plus, per row, the sum of (a_ij - fl32(a_ij)) over the row's
// off-diagonals, so every row of the inner operator sums to the same value
// as A's row (to FP64 rounding).  The rounding error is then
// E = sum_{i<j} e_ij (u_i - u_j)(u_i - u_j)^T with |e_ij| <= 2^-24 |a_ij|,
// and for a graph Laplacian plus a non-negative diagonal shift
// |x^T E x| <= 2^-24 x^T A x: the inner operator stays SPD and within a
// factor 1 +- 6e-8 of A however small the shift.  Plain rounding moves a
// row sum by up to ~2^-24 sqrt(deg) |a_i|, which for a power-law graph's
// hub rows (degree ~1e4) exceeds a 1e-6 relative diagonal shift and makes
// the FP32 operator indefinite.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "lap_core.hpp"
#include "spmv_core.hpp"

namespace lap {

namespace {

struct MixCtl {
  double rho;        // r32.w32 of the iteration in flight
  double rho_prev;   // generic path: previous rho
  double tol2res0;   // (tol * ||r0||)^2
  double maxr2;      // squared FP64 residual at the last refinement
  double delta2;     // refinement trigger (squared)
  double res0, res;
  double pend_rho, pend_maxr2;
  int iter, maxit, done, flag;
  int pend_iter, pend_valid, last_iter, refinements;
  int pend_refinements, pad;
};

constexpr int kQ0 = 0, kQ1 = kMaxPartialBlocks, kQ2 = 2 * kMaxPartialBlocks,
              kQ3 = 3 * kMaxPartialBlocks;

// Refinement decision of iteration `it`: a pure function of the recursive
// residual and committed fields, evaluated identically by every kernel.
__device__ __forceinline__ bool need_refine(double rr, int it, const MixCtl *c) {
  return !(rr >= c->delta2 * c->maxr2) || rr < c->tol2res0 || it > c->maxit;
}

// --------------------------------------------------------------- kernels

__global__ __launch_bounds__(kBlock) void k_to_f32(int n, const double *__restrict__ s,
                                                   float *__restrict__ d) {
  const size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < (size_t)n; i += stride)
    d[i] = (float)s[i];
}

template <class VT>
__global__ __launch_bounds__(kBlock) void k_zero_v(int n, VT *__restrict__ d) {
  const size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < (size_t)n; i += stride)
    d[i] = VT(0);
}

// FP64 residual epilogue: r = (VT)(b - A x); sums of r.r (FP64) and, with
// Jacobi, r.w and p = w (first direction) for w = d_r r.
template <class VT>
struct Resid64Epi {
  const double *rhs;
  VT *r32;
  const VT *d_r;  // null: no Jacobi terms
  VT *p_init;     // null: do not set p
  double rr, rz;
  __device__ double init(int) const { return 0.0; }
  __device__ void store(int row, double s) {
    const double v = -(s - rhs[row]);
    const VT rf = (VT)v;
    r32[row] = rf;
    rr += v * v;
    if (d_r) {
      const VT wf = d_r[row] * rf;
      rz += (double)rf * (double)wf;
      if (p_init) p_init[row] = wf;
    }
  }
};

// r = b - A x (FP64), initial version: always runs.
template <int BS, int TILE, class VT>
__global__ __launch_bounds__(BS) void k_mx_init(int nblk, const int *__restrict__ rowblk,
                                                const int *__restrict__ ia,
                                                const int *__restrict__ ja,
                                                const double *__restrict__ a,
                                                const double *__restrict__ x,
                                                const double *__restrict__ rhs, VT *r32,
                                                const VT *d_r, VT *p32, double *part,
                                                int walk) {
  __shared__ double prod[TILE];
  __shared__ double red[BS / kWave];
  Resid64Epi<VT> epi{rhs, r32, d_r, p32, 0.0, 0.0};
  spmv_walk<BS, TILE>(block_walk(nblk, walk), rowblk, ia, ja, a, x, 1.0, prod, red, epi);
  const double t_rz = block_sum<BS>(epi.rz, red);
  const double t_rr = block_sum<BS>(epi.rr, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = t_rz;
    part[2 * blockIdx.x + 1] = t_rr;
  }
}

// one workgroup: ||r0||, tolerance, initial rho
template <int BS>
__global__ __launch_bounds__(BS) void k_mx_start(const double *part, int cnt, double tol,
                                                 double delta, MixCtl *ctl, double *hist) {
  __shared__ double red[BS / kWave];
  const double rz = sum_partials<BS>(part, cnt, 2, red);
  const double rr = sum_partials<BS>(part + 1, cnt, 2, red);
  if (threadIdx.x == 0) {
    const double res0 = sqrt(rr);
    hist[0] = res0;
    ctl->res0 = res0;
    ctl->res = res0;
    ctl->tol2res0 = (tol * res0) * (tol * res0);
    ctl->maxr2 = rr;
    ctl->delta2 = delta * delta;
    ctl->rho = rz;
    if (!isfinite(res0)) {
      ctl->flag = 2;
      ctl->done = 1;
    } else if (res0 == 0.0) {
      ctl->done = 1;
    }
  }
}

template <class VT>
struct PqMxEpi {
  VT *q;
  const VT *p;
  const double *diag;
  double acc;
  __device__ double init(int row) const { return diag[row] * (double)p[row]; }
  __device__ void store(int row, double s) {
    const VT qf = (VT)s;
    st_nt(q + row, qf);  // q is read once, by the next kernel
    acc += (double)p[row] * (double)qf;
  }
};

// q = A32 p (FP32 off-diagonals + FP64 diagonal, FP64 row sums); partials
// of p.q.  Workgroup 0 commits the pending control fields.
template <int BS, int TILE, class VT>
__global__ __launch_bounds__(BS) void k_mx_spmv_pq(int nblk, const int *__restrict__ rowblk,
                                                   const int *__restrict__ ia,
                                                   const int *__restrict__ ja,
                                                   const float *__restrict__ a,
                                                   const double *__restrict__ diag,
                                                   const VT *__restrict__ p,
                                                   VT *__restrict__ q, MixCtl *ctl,
                                                   double *part, int walk) {
  if (ctl->done) return;
  if (blockIdx.x == 0 && threadIdx.x == 0 && ctl->pend_valid) {
    ctl->iter = ctl->pend_iter;
    ctl->rho = ctl->pend_rho;
    ctl->maxr2 = ctl->pend_maxr2;
    ctl->refinements = ctl->pend_refinements;
    ctl->pend_valid = 0;
  }
  __shared__ double prod[TILE];
  __shared__ double red[BS / kWave];
  PqMxEpi<VT> epi{q, p, diag, 0.0};
  const BlockWalk w = block_walk(nblk, walk);
  for (int s = w.begin; s < w.end; s += w.step) {
    const int blk = w.block(s);
    if (blk >= nblk) continue;
    spmv_block_rows<BS, TILE, PqMxEpi<VT>, float, VT>(blk, rowblk, ia, ja, a, p, 1.0, prod, red,
                                                      epi);
  }
  const double t = block_sum<BS>(epi.acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// The same over the value-indexed copy: off-diagonal value = dict[vi[e]],
// the table staged in LDS; bitwise the same products and sums.
template <int BS, int TILE, class VT>
__global__ __launch_bounds__(BS) void k_mx_spmv_pq_vi(int nblk, const int *__restrict__ rowblk,
                                                      const int *__restrict__ ia,
                                                      const int *__restrict__ ja,
                                                      const unsigned char *__restrict__ vi,
                                                      const float *__restrict__ dict, int nd,
                                                      const double *__restrict__ diag,
                                                      const VT *__restrict__ p,
                                                      VT *__restrict__ q, MixCtl *ctl,
                                                      double *part, int walk) {
  if (ctl->done) return;
  if (blockIdx.x == 0 && threadIdx.x == 0 && ctl->pend_valid) {
    ctl->iter = ctl->pend_iter;
    ctl->rho = ctl->pend_rho;
    ctl->maxr2 = ctl->pend_maxr2;
    ctl->refinements = ctl->pend_refinements;
    ctl->pend_valid = 0;
  }
  __shared__ double prod[TILE];
  __shared__ double red[BS / kWave];
  __shared__ float table[256];
  for (int i = threadIdx.x; i < nd; i += BS) table[i] = dict[i];
  __syncthreads();
  PqMxEpi<VT> epi{q, p, diag, 0.0};
  const BlockWalk w = block_walk(nblk, walk);
  for (int s = w.begin; s < w.end; s += w.step) {
    const int blk = w.block(s);
    if (blk >= nblk) continue;
    spmv_block_rows_vi<BS, TILE, PqMxEpi<VT>, float, VT>(blk, rowblk, ia, ja, vi, table, p, 1.0,
                                                        prod, red, epi);
  }
  const double t = block_sum<BS>(epi.acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// The same over the delta-indexed copy: code = v | d << 8, value dict[v],
// column row + ddict[d] (tables in LDS); the block's row pointers and codes
// are staged into LDS with coalesced loads, then one lane per row (rows of at
// most 64 entries) sums its products in CSR order after the diagonal term:
// bitwise the CSR kernel's rows.
template <int BS, class VT>
__global__ __launch_bounds__(BS) void k_mx_spmv_pq_dvi(int nblk, const int *__restrict__ rowblk,
                                                       const int *__restrict__ ia,
                                                       const unsigned short *__restrict__ code,
                                                       const float *__restrict__ dict, int nd,
                                                       const int *__restrict__ ddict, int ndl,
                                                       const double *__restrict__ diag,
                                                       const VT *__restrict__ p,
                                                       VT *__restrict__ q, MixCtl *ctl,
                                                       double *part, int walk) {
  if (ctl->done) return;
  if (blockIdx.x == 0 && threadIdx.x == 0 && ctl->pend_valid) {
    ctl->iter = ctl->pend_iter;
    ctl->rho = ctl->pend_rho;
    ctl->maxr2 = ctl->pend_maxr2;
    ctl->refinements = ctl->pend_refinements;
    ctl->pend_valid = 0;
  }
  __shared__ double red[BS / kWave];
  __shared__ float vt[256];
  __shared__ int dt[256];
  __shared__ unsigned short sc[kSpmvTile];
  __shared__ int rp[2 * BS + 1];
  for (int i = threadIdx.x; i < nd; i += BS) vt[i] = dict[i];
  for (int i = threadIdx.x; i < ndl; i += BS) dt[i] = ddict[i];
  PqMxEpi<VT> epi{q, p, diag, 0.0};
  const BlockWalk w = block_walk(nblk, walk);
  for (int s = w.begin; s < w.end; s += w.step) {
    const int b = w.block(s);
    if (b >= nblk) continue;
    const int r0 = rowblk[b], r1 = rowblk[b + 1];
    const int e0 = ia[r0];
    for (int i = threadIdx.x; i <= r1 - r0; i += BS) rp[i] = ia[r0 + i] - e0;
    __syncthreads();
    const int cnt = rp[r1 - r0];
    for (int k = threadIdx.x; k < cnt; k += BS) sc[k] = code[e0 + k];
    __syncthreads();
    for (int row = r0 + (int)threadIdx.x; row < r1; row += BS) {
      const int k1 = rp[row - r0 + 1];
      double sum = epi.init(row);
      for (int k = rp[row - r0]; k < k1; k += 8) {
        unsigned c[8];
        double xv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) c[j] = k + j < k1 ? (unsigned)sc[k + j] : 0u;
#pragma unroll
        for (int j = 0; j < 8; ++j)
          xv[j] = k + j < k1 ? (double)p[row + dt[c[j] >> 8]] : 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (k + j < k1) sum += (1.0 * (double)vt[c[j] & 255u]) * xv[j];
      }
      epi.store(row, sum);
    }
    __syncthreads();
  }
  const double t = block_sum<BS>(epi.acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// alpha = rho/(p.q); x32 += alpha p; r32 -= alpha q (each rounded once from
// FP64 to VT); partials of r.r and, with Jacobi, r.w for w = d_r r (stride
// 2: [rz, rr]).
template <int BS, bool JAC, class VT>
__global__ __launch_bounds__(BS) void k_mx_update(int n, VT *__restrict__ x32,
                                                  VT *__restrict__ r32,
                                                  const VT *__restrict__ p,
                                                  const VT *__restrict__ q,
                                                  const VT *__restrict__ d_r,
                                                  const MixCtl *ctl, const double *pq_part,
                                                  int pq_cnt, double *part) {
  if (ctl->done) return;
  __shared__ double red[BS / kWave];
  const double pq = sum_partials<BS>(pq_part, pq_cnt, 1, red);
  const double alpha = ctl->rho / pq;
  double rz = 0.0, rr = 0.0;
  const size_t stride = (size_t)gridDim.x * BS;
  // streamed vectors: non-temporal loads and stores (values unchanged)
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride) {
    st_nt(x32 + i, (VT)((double)ld_nt(x32 + i) + alpha * (double)ld_nt(p + i)));
    const VT rf = (VT)((double)ld_nt(r32 + i) - alpha * (double)ld_nt(q + i));
    st_nt(r32 + i, rf);
    rr += (double)rf * (double)rf;
    if (JAC) {
      const VT wf = ld_nt(d_r + i) * rf;
      rz += (double)rf * (double)wf;
    }
  }
  const double t_rz = block_sum<BS>(rz, red);
  const double t_rr = block_sum<BS>(rr, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = t_rz;
    part[2 * blockIdx.x + 1] = t_rr;
  }
}

// refinement, step 1 of the generic loop: x += x32, x32 = 0 (only when this
// iteration refines); with `xsum` given, partial sums of the new x (for the
// deflation shift).  The Jacobi loop folds in k_mx_jac_p instead.
template <int BS, class VT>
__global__ __launch_bounds__(BS) void k_mx_fold(int n, double *__restrict__ x,
                                                VT *__restrict__ x32, const MixCtl *ctl,
                                                const double *upd_part, int upd_cnt,
                                                double *xsum) {
  if (ctl->done) return;
  __shared__ double red[BS / kWave];
  const double rr = sum_partials<BS>(upd_part + 1, upd_cnt, 2, red);
  if (!need_refine(rr, ctl->iter + 1, ctl)) return;
  double s = 0.0;
  const size_t stride = (size_t)gridDim.x * BS;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride) {
    const double v = x[i] + (double)x32[i];
    x[i] = v;
    x32[i] = VT(0);
    s += v;
  }
  if (xsum) {
    const double t = block_sum<BS>(s, red);
    if (threadIdx.x == 0) xsum[blockIdx.x] = t;
  }
}

// refinement, step 2: r32 = (float)(b - A x) in FP64; partials [rz, rr].
// With x32 given (the Jacobi loop, whose fold comes after this kernel, in
// k_mx_jac_p), the residual is of x + x32: the folded x, bit for bit.
template <int BS, int TILE, class VT>
__global__ __launch_bounds__(BS) void k_mx_resid(int nblk, const int *__restrict__ rowblk,
                                                 const int *__restrict__ ia,
                                                 const int *__restrict__ ja,
                                                 const double *__restrict__ a,
                                                 const double *__restrict__ x,
                                                 const VT *__restrict__ x32,
                                                 const double *__restrict__ rhs, VT *r32,
                                                 const VT *d_r, const MixCtl *ctl,
                                                 const double *upd_part, int upd_cnt,
                                                 double *part, int walk) {
  if (ctl->done) return;
  __shared__ double prod[TILE];
  __shared__ double red[BS / kWave];
  const double rr_rec = sum_partials<BS>(upd_part + 1, upd_cnt, 2, red);
  if (!need_refine(rr_rec, ctl->iter + 1, ctl)) return;
  Resid64Epi<VT> epi{rhs, r32, d_r, nullptr, 0.0, 0.0};
  if (x32)
    spmv_walk_sum<BS, TILE>(block_walk(nblk, walk), rowblk, ia, ja, a, x, x32, 1.0, prod, red,
                            epi);
  else
    spmv_walk<BS, TILE>(block_walk(nblk, walk), rowblk, ia, ja, a, x, 1.0, prod, red, epi);
  const double t_rz = block_sum<BS>(epi.rz, red);
  const double t_rr = block_sum<BS>(epi.rr, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = t_rz;
    part[2 * blockIdx.x + 1] = t_rr;
  }
}

// Outcome of an iteration from the update and refinement partials.
struct MxStep {
  bool refine;
  double rz, rr, res;
  int it, flag;  // flag -1: continue
};

template <int BS>
__device__ __forceinline__ MxStep mx_step(const MixCtl *ctl, const double *upd_part, int upd_cnt,
                                          const double *ref_part, int ref_cnt, double *red) {
  MxStep s;
  const double rz_u = sum_partials<BS>(upd_part, upd_cnt, 2, red);
  const double rr_u = sum_partials<BS>(upd_part + 1, upd_cnt, 2, red);
  s.it = ctl->iter + 1;
  s.refine = need_refine(rr_u, s.it, ctl);
  s.rz = rz_u;
  s.rr = rr_u;
  if (s.refine) {
    s.rz = sum_partials<BS>(ref_part, ref_cnt, 2, red);
    s.rr = sum_partials<BS>(ref_part + 1, ref_cnt, 2, red);
  }
  s.res = sqrt(s.rr);
  s.flag = -1;
  if (s.refine) {  // decisions only on an FP64 residual
    if (!isfinite(s.res)) s.flag = 2;
    else if (s.rr < ctl->tol2res0) s.flag = 0;
    else if (s.it > ctl->maxit) s.flag = 1;
  }
  return s;
}

// Jacobi: convergence bookkeeping, the refinement's fold (x += x32, x32 = 0,
// after k_mx_resid has used x + x32), beta, p = beta p + d_r r
template <int BS, class VT>
__global__ __launch_bounds__(BS) void k_mx_jac_p(int n, const VT *__restrict__ r32,
                                                 const VT *__restrict__ d_r,
                                                 VT *__restrict__ p, double *__restrict__ x,
                                                 VT *__restrict__ x32, MixCtl *ctl,
                                                 double *hist, const double *upd_part,
                                                 int upd_cnt, const double *ref_part,
                                                 int ref_cnt) {
  if (ctl->done) return;
  __shared__ double red[BS / kWave];
  const MxStep s = mx_step<BS>(ctl, upd_part, upd_cnt, ref_part, ref_cnt, red);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    hist[s.it] = s.res;
    ctl->res = s.res;
    ctl->last_iter = s.it;
    if (s.flag >= 0) {
      ctl->flag = s.flag;
      ctl->refinements = ctl->refinements + 1;
      ctl->done = 1;
    } else {
      ctl->pend_iter = s.it;
      ctl->pend_rho = s.rz;
      ctl->pend_maxr2 = s.refine ? s.rr : ctl->maxr2;
      ctl->pend_refinements = ctl->refinements + (s.refine ? 1 : 0);
      ctl->pend_valid = 1;
    }
  }
  const size_t stride = (size_t)gridDim.x * BS;
  if (s.refine)  // also on the last iteration: x is the solution
    for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride) {
      x[i] = x[i] + (double)x32[i];
      x32[i] = VT(0);
    }
  if (s.flag >= 0) return;
  const double beta = s.rz / ctl->rho;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride) {
    const VT wf = ld_nt(d_r + i) * ld_nt(r32 + i);
    st_nt(p + i, (VT)((double)ld_nt(p + i) * beta + (double)wf));
  }
}

// generic: partials of r.w; deflated (DEFL): [r.w, sum w, sum r] per
// workgroup
template <int BS, class VT, bool DEFL>
__global__ __launch_bounds__(BS) void k_mx_rw(int n, const VT *__restrict__ r,
                                              const VT *__restrict__ w, const MixCtl *ctl,
                                              double *part) {
  if (ctl->done) return;
  __shared__ double red[BS / kWave];
  double s = 0.0, sw = 0.0, sr = 0.0;
  const size_t stride = (size_t)gridDim.x * BS;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride) {
    s += (double)r[i] * (double)w[i];
    if (DEFL) {
      sw += (double)w[i];
      sr += (double)r[i];
    }
  }
  const double t = block_sum<BS>(s, red);
  if (!DEFL) {
    if (threadIdx.x == 0) part[blockIdx.x] = t;
    return;
  }
  const double tw = block_sum<BS>(sw, red);
  const double tr = block_sum<BS>(sr, red);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = t;
    part[3 * blockIdx.x + 1] = tw;
    part[3 * blockIdx.x + 2] = tr;
  }
}

// generic: rho, p = w (first) or beta p + w.  Deflated (DEFL): w is taken
// with its mean removed (w - mu, mu = sum w / n), so p, q = A p and r stay
// orthogonal to the constant vector, and rho = r.(w - mu) = r.w - mu sum r.
template <int BS, class VT, bool DEFL>
__global__ __launch_bounds__(BS) void k_mx_pupd(int n, const VT *__restrict__ w,
                                                VT *__restrict__ p, MixCtl *ctl,
                                                const double *part, int cnt) {
  if (ctl->done) return;
  __shared__ double red[BS / kWave];
  double rho, mu = 0.0;
  if (DEFL) {
    const double rw = sum_partials<BS>(part, cnt, 3, red);
    const double sw = sum_partials<BS>(part + 1, cnt, 3, red);
    const double sr = sum_partials<BS>(part + 2, cnt, 3, red);
    mu = sw / (double)n;
    rho = rw - mu * sr;
  } else {
    rho = sum_partials<BS>(part, cnt, 1, red);
  }
  const bool first = ctl->iter == 0;
  const double beta = first ? 0.0 : rho / ctl->rho_prev;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->rho = rho;
  const size_t stride = (size_t)gridDim.x * BS;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride) {
    const double wi = DEFL ? (double)w[i] - mu : (double)w[i];
    p[i] = first ? (VT)wi : (VT)((double)p[i] * beta + wi);
  }
}

// Deflation of the constant vector, for matrices whose rows all sum to the
// same s > 0 (a graph Laplacian plus s I): A 1 = s 1, so the solution's
// constant component is mean(b) / s and the rest lies in 1-perp, where the
// inner iteration is kept.  x's mean is set to `target` = mean(b) / s
// before the first residual and at every refinement.
// Partial sums of the caller's initial x.
template <int BS>
__global__ __launch_bounds__(BS) void k_mx_xsum(int n, const double *__restrict__ x,
                                                double *part) {
  __shared__ double red[BS / kWave];
  double s = 0.0;
  const size_t stride = (size_t)gridDim.x * BS;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride) s += x[i];
  const double t = block_sum<BS>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// x += target - mean(x): before the first residual, and after the fold on
// refinement iterations (`upd_part` given).
template <int BS>
__global__ __launch_bounds__(BS) void k_mx_xshift(int n, double *__restrict__ x, double target,
                                                  const MixCtl *ctl, const double *upd_part,
                                                  int upd_cnt, const double *part, int cnt) {
  if (ctl->done) return;
  __shared__ double red[BS / kWave];
  if (upd_part) {
    const double rr = sum_partials<BS>(upd_part + 1, upd_cnt, 2, red);
    if (!need_refine(rr, ctl->iter + 1, ctl)) return;
  }
  const double d = target - sum_partials<BS>(part, cnt, 1, red) / (double)n;
  const size_t stride = (size_t)gridDim.x * BS;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride) x[i] += d;
}

// Maximum over a block of BS threads; the result is valid in thread 0.
template <int BS>
__device__ __forceinline__ double block_max(double v, double *lds) {
  for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o));
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double r = -INFINITY;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < BS / kWave; ++i) r = fmax(r, lds[i]);
  }
  __syncthreads();
  return r;
}

// Row statistics of A for the deflation test: per workgroup the smallest
// and largest row sum (CSR order), the largest diagonal and the sum of b.
template <int BS>
__global__ __launch_bounds__(BS) void k_mx_rowstats(int n, const int *__restrict__ ia,
                                                    const int *__restrict__ ja,
                                                    const double *__restrict__ a,
                                                    const double *__restrict__ b,
                                                    double *part) {
  __shared__ double red[BS / kWave];
  double lo = INFINITY, hi = -INFINITY, dmax = 0.0, sb = 0.0;
  const size_t stride = (size_t)gridDim.x * BS;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride) {
    double rs = 0.0;
    for (int e = ia[i]; e < ia[i + 1]; ++e) {
      rs += a[e];
      if (ja[e] == (int)i) dmax = fmax(dmax, a[e]);
    }
    lo = fmin(lo, rs);
    hi = fmax(hi, rs);
    sb += b[i];
  }
  const double t_lo = -block_max<BS>(-lo, red);
  const double t_hi = block_max<BS>(hi, red);
  const double t_d = block_max<BS>(dmax, red);
  const double t_b = block_sum<BS>(sb, red);
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x] = t_lo;
    part[4 * blockIdx.x + 1] = t_hi;
    part[4 * blockIdx.x + 2] = t_d;
    part[4 * blockIdx.x + 3] = t_b;
  }
}

// generic, one workgroup: convergence bookkeeping
template <int BS>
__global__ __launch_bounds__(BS) void k_mx_conv(MixCtl *ctl, double *hist,
                                                const double *upd_part, int upd_cnt,
                                                const double *ref_part, int ref_cnt) {
  if (ctl->done) return;
  __shared__ double red[BS / kWave];
  const MxStep s = mx_step<BS>(ctl, upd_part, upd_cnt, ref_part, ref_cnt, red);
  if (threadIdx.x == 0) {
    hist[s.it] = s.res;
    ctl->res = s.res;
    ctl->last_iter = s.it;
    ctl->iter = s.it;
    ctl->rho_prev = ctl->rho;
    if (s.refine) {
      ctl->maxr2 = s.rr;
      ctl->refinements = ctl->refinements + 1;
    }
    if (s.flag >= 0) {
      ctl->flag = s.flag;
      ctl->done = 1;
    }
  }
}

template <class VT>
__global__ __launch_bounds__(kBlock) void k_copy_v(int n, const VT *__restrict__ s,
                                                   VT *__restrict__ d,
                                                     const int *__restrict__ done) {
  if (done && *done) return;
  const size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < (size_t)n; i += stride)
    d[i] = s[i];
}

// ---------------------------------------------------------------- plans

// Off-diagonal pattern of A: entries per row without the diagonal.
__global__ __launch_bounds__(kBlock) void k_mx_count(int n, const int *__restrict__ ia,
                                                     const int *__restrict__ ja,
                                                     int *__restrict__ cnt) {
  for (int row = blockIdx.x * kBlock + threadIdx.x; row < n; row += gridDim.x * kBlock) {
    int c = 0;
    for (int e = ia[row]; e < ia[row + 1]; ++e) c += ja[e] != row;
    cnt[row] = c;
  }
}

__global__ __launch_bounds__(kBlock) void k_mx_pattern(int n, const int *__restrict__ ia,
                                                       const int *__restrict__ ja,
                                                       const int *__restrict__ io,
                                                       int *__restrict__ jo) {
  for (int row = blockIdx.x * kBlock + threadIdx.x; row < n; row += gridDim.x * kBlock) {
    int pos = io[row];
    for (int e = ia[row]; e < ia[row + 1]; ++e)
      if (ja[e] != row) jo[pos++] = ja[e];
  }
}

// The inner operator's values from A's, every solve: FP32 off-diagonals (in
// the off-diagonal pattern's order; a64 gets the same entries in FP64 when
// not null) and the FP64 diagonal plus what rounding took from the row's
// off-diagonals, accumulated in CSR order.
__global__ __launch_bounds__(kBlock) void k_mx_values(int n, const int *__restrict__ ia,
                                                      const int *__restrict__ ja,
                                                      const double *__restrict__ a,
                                                      const int *__restrict__ io,
                                                      float *__restrict__ ao,
                                                      double *__restrict__ dg,
                                                      double *__restrict__ a64) {
  for (int row = blockIdx.x * kBlock + threadIdx.x; row < n; row += gridDim.x * kBlock) {
    double d = 0.0, comp = 0.0;
    int pos = io[row];
    for (int e = ia[row]; e < ia[row + 1]; ++e) {
      const double v = a[e];
      if (ja[e] == row) {
        d += v;
      } else {
        const float f = (float)v;
        ao[pos] = f;
        if (a64) a64[pos] = v;
        ++pos;
        comp += v - (double)f;
      }
    }
    dg[row] = d + comp;
  }
}

// Per sparsity pattern (ia, ja, n, nnz): the off-diagonal pattern, built on
// the device once.  Everything derived from values (the FP32 off-diagonals,
// the compensated diagonal, the value- and delta-indexed copies, the FP32
// 1/diag and IC0 factor) is rebuilt on the device at every solve, since the
// caller may change A, d_r or the factor in place between solves.
struct MixPlan {
  int n = 0, nnz_off = 0;
  DevBufPtr ia_o, ja_o;  // off-diagonal pattern
  DevBufPtr a32_o, diag; // this solve's values
  // value-indexed copy (spmv_format >= 1, at most 255 distinct off-diagonal
  // values): one byte per entry and the FP32 table; the delta-indexed copy
  // (spmv_format 2, also at most 255 distinct col - row and rows of at most
  // 64 entries): code = value index | offset index << 8
  DevBufPtr a64_o, vi_o, dict64, dict_o, code_o, ddict_o;
  int nd = 0, ndl = 0;
  DevBufPtr dr32, l32, u32;  // this solve's FP32 1/diag and IC0 factor
  TriChunks32 lch, uch;      // the factor's chunked-solve records
  const SpmvPlan *off = nullptr;
};
using MixKey = std::tuple<const int *, const int *, int, int>;
std::map<MixKey, std::unique_ptr<MixPlan>> g_mix;

DevBufPtr &sized(DevBufPtr &b, size_t bytes) {
  if (!b || b->bytes < bytes) b = dev_alloc(bytes);
  return b;
}

MixPlan *mix_plan(int n, int nnz, const int *ia, const int *ja) {
  MixKey key{ia, ja, n, nnz};
  auto it = g_mix.find(key);
  if (it != g_mix.end()) return it->second.get();
  auto P = std::make_unique<MixPlan>();
  P->n = n;
  P->ia_o = dev_alloc(sizeof(int) * ((size_t)n + 1));
  DevBuf cnt(sizeof(int) * ((size_t)n + 1)), ws(scan_ws_bytes(n) + 16);
  const unsigned g = grid_for((size_t)n, 1, kBlock, 65536);
  k_mx_count<<<g, kBlock, 0, stream()>>>(n, ia, ja, cnt.as<int>());
  const long long total = scan_exclusive(cnt.as<int>(), n, P->ia_o->as<int>(),
                                         ws.as<long long>());
  if (total < 0) return nullptr;
  P->nnz_off = (int)total;
  P->ja_o = dev_alloc(sizeof(int) * (size_t)std::max<long long>(total, 1));
  k_mx_pattern<<<g, kBlock, 0, stream()>>>(n, ia, ja, P->ia_o->as<int>(), P->ja_o->as<int>());
  P->a32_o = dev_alloc(sizeof(float) * (size_t)std::max<long long>(total, 1));
  P->diag = dev_alloc(sizeof(double) * (size_t)n);
  sync();
  P->off = spmv_plan(n, P->nnz_off, P->ia_o->as<int>(), P->ja_o->as<int>());
  if (!P->off || lap_last_status() == LAP_ERR_HIP) return nullptr;
  MixPlan *raw = P.get();
  g_mix[key] = std::move(P);
  return raw;
}

// This solve's inner-operator values (and indexed copies when `format` asks).
bool mix_values(MixPlan &M, const int *ia, const int *ja, const double *a, int format) {
  const int n = M.n;
  const size_t no = (size_t)std::max(M.nnz_off, 1);
  double *a64 = nullptr;
  if (format >= 1) a64 = sized(M.a64_o, sizeof(double) * no)->as<double>();
  k_mx_values<<<grid_for((size_t)n, 1, kBlock, 65536), kBlock, 0, stream()>>>(
      n, ia, ja, a, M.ia_o->as<int>(), M.a32_o->as<float>(), M.diag->as<double>(), a64);
  M.nd = M.ndl = 0;
  if (format >= 1 && M.nnz_off > 0) {
    sized(M.dict64, sizeof(double) * 256);
    sized(M.dict_o, sizeof(float) * 256);
    int nv = 0;
    if (format == 2) {
      sized(M.code_o, sizeof(unsigned short) * no);
      sized(M.ddict_o, sizeof(int) * 256);
      int ndl = 0;
      const int r = dvi_encode(n, M.nnz_off, M.ia_o->as<int>(), M.ja_o->as<int>(), a64, 64,
                               M.code_o->as<unsigned short>(), M.dict64->as<double>(),
                               M.ddict_o->as<int>(), &nv, &ndl);
      if (r < 0) return false;
      if (r == 1) M.ndl = ndl;
    }
    if (M.ndl == 0) {  // format 1, or 2 falling back
      sized(M.vi_o, no);
      nv = vi_encode(M.nnz_off, a64, M.vi_o->as<unsigned char>(), M.dict64->as<double>());
      if (nv < 0) return false;
    }
    // the table in FP32: fl32 of each distinct double (fl32 may map two of
    // them to one float; each index still gives its entry's FP32 value)
    if (nv > 0)
      k_to_f32<<<1, kBlock, 0, stream()>>>(nv, M.dict64->as<double>(), M.dict_o->as<float>());
    M.nd = nv;
  }
  return lap_last_status() != LAP_ERR_HIP;
}

float *to_f32(DevBufPtr &buf, const double *src, int n) {
  sized(buf, sizeof(float) * (size_t)std::max(n, 1));
  if (n > 0)
    k_to_f32<<<grid_for((size_t)n, 4, kBlock, 2048), kBlock, 0, stream()>>>(n, src,
                                                                         buf->as<float>());
  return buf->as<float>();
}

}  // namespace

void release_mix_plans() { g_mix.clear(); }
void forget_mix_plans(const void *p) {
  for (auto it = g_mix.begin(); it != g_mix.end();) {
    const MixKey &k = it->first;
    if (std::get<0>(k) == p || std::get<1>(k) == p)
      it = g_mix.erase(it);
    else
      ++it;
  }
}

struct MixOut {
  int iterations = 0, flag = 2, refinements = 0;
  double res0 = 0, res = 0, gpu_ms = 0;
};

// 1/diag for the inner Jacobi step in the inner vectors' type
template <class VT>
const VT *inner_d_r(MixPlan *M, const double *d_r, int n);
template <>
const float *inner_d_r<float>(MixPlan *M, const double *d_r, int n) {
  return to_f32(M->dr32, d_r, n);
}
template <>
const double *inner_d_r<double>(MixPlan *, const double *d_r, int) {
  return d_r;
}

// The inner IC0 step.  FP32 vectors: the FP32 copy of the factor.  FP64
// vectors: the FP64 factor itself -- the level-scheduled solves are bound by
// latency, not bytes, so FP32 factor values would save no time and their
// rounding costs iterations on badly conditioned systems.
inline void ic_inner(const IcPlan &P, const MixPlan &M, const float *l32, const float *u32,
                     const float *x, float *y, float *aux, const int *done) {
  tri_solve32_launch(*P.lplan, P.lja->as<int>(), l32, x, aux, done, &M.lch);
  tri_solve32_launch(*P.uplan, P.uja->as<int>(), u32, aux, y, done, &M.uch);
}
inline void ic_inner(const IcPlan &P, const MixPlan &, const float *, const float *,
                     const double *x, double *y, double *aux, const int *done) {
  ic_apply_launch(P, x, y, aux, done);
}

template <class VT>
static int run_mixed(int n, int nnz, const int *ia, const int *ja, const double *a, double *x,
                     const double *b, double tol, lap_pdata *pd, int maxit, double delta,
                     double *h_hist, MixOut &out) {
  if (n <= 0 || nnz < 0 || maxit < 0 || !(delta > 0.0 && delta < 1.0)) {
    set_error(LAP_ERR_ARG, "pcg_mixed: bad arguments n=%d maxit=%d delta=%g", n, maxit, delta);
    return LAP_ERR_ARG;
  }
  for (const void *ptr : {(const void *)ia, (const void *)ja, (const void *)a, (const void *)x,
                          (const void *)b})
    if (!is_device_ptr(ptr)) {
      set_error(LAP_ERR_ARG, "pcg_mixed: argument %p is not device memory", ptr);
      return LAP_ERR_ARG;
    }
  const char *op = pd && pd->prec_op ? pd->prec_op : "";
  const bool jac = !strcmp(op, "line_jacobi");
  const bool ic = !strcmp(op, "ichol");
  const bool none = !strcmp(op, "none");
  if (!jac && !ic && !none) {
    set_error(LAP_ERR_PREC, "pcg_mixed: preconditioner '%s' not supported (line_jacobi, ichol, "
              "none)", op);
    return LAP_ERR_PREC;
  }
  const SpmvPlan *A64 = spmv_plan(n, nnz, ia, ja);
  if (!A64) return lap_last_status();
  MixPlan *M = mix_plan(n, nnz, ia, ja);
  if (!M || !mix_values(*M, ia, ja, a, options().spmv_format)) return lap_last_status();
  const VT *d_r32 = nullptr;
  if (jac) {
    if (!is_device_ptr(pd->d_r)) {
      set_error(LAP_ERR_ARG, "pcg_mixed: prec_data->d_r is not device memory");
      return LAP_ERR_ARG;
    }
    d_r32 = inner_d_r<VT>(M, pd->d_r, n);
  }
  const IcPlan *icp = nullptr;
  float *l32 = nullptr, *u32 = nullptr;
  if (ic) {
    icp = ic_plan_find(pd->ichol_vals);
    if (!icp) {
      set_error(LAP_ERR_PREC, "pcg_mixed: no IC0 factor for ichol_vals (call initialize_ichol)");
      return LAP_ERR_PREC;
    }
    if (sizeof(VT) == sizeof(float)) {  // this solve's FP32 copy of the factor
      l32 = to_f32(M->l32, icp->lval->as<double>(), icp->nnzL);
      u32 = to_f32(M->u32, icp->uval->as<double>(), icp->nnzL);
      tri_ch_pack32(*icp->lplan, icp->lja->as<int>(), l32, M->lch);
      tri_ch_pack32(*icp->uplan, icp->uja->as<int>(), u32, M->uch);
    }
  }

  auto fvec = [&](const char *tag) {
    return reinterpret_cast<VT *>(scratch_vec(tag, ((size_t)n * sizeof(VT) + 7) / 8 + 1));
  };
  VT *r32 = fvec("mx_r"), *p32 = fvec("mx_p"), *q32 = fvec("mx_q"), *x32 = fvec("mx_x"),
        *w32 = fvec("mx_w"), *aux32 = fvec("mx_aux");
  MixCtl *ctl = reinterpret_cast<MixCtl *>(scratch_vec("mx_ctl", sizeof(MixCtl) / 8 + 1));
  double *hist = scratch_vec("mx_hist", (size_t)maxit + 2);
  const int *done = &ctl->done;
  ReduceWs &ws = reduce_ws();
  double *Q0 = ws.partials + kQ0, *Q1 = ws.partials + kQ1, *Q2 = ws.partials + kQ2,
         *Q3 = ws.partials + kQ3;
  const unsigned gA = spmv_grid(*A64), gO = spmv_grid(*M->off);
  // FP32 vectors carry half the bytes per element: their kernels run best with
  // twice the FP64 kernels' workgroups (measured at 256^3: 1536 against 768)
  // (capped so that the two partials per workgroup of k_mx_update stay
  // inside their kMaxPartialBlocks region)
  const unsigned gV = grid_for(
      (size_t)n, 4, kBlock,
      std::min<unsigned>((unsigned)options().vec_grid * (sizeof(VT) == 4 ? 2u : 1u),
                         kMaxPartialBlocks / 2));
  if (2 * gA > (unsigned)kMaxPartialBlocks || gO > (unsigned)kMaxPartialBlocks) {
    set_error(LAP_ERR_ARG, "pcg_mixed: SpMV grids %u/%u exceed the partial-sum regions", gA, gO);
    return LAP_ERR_ARG;
  }
  const int walk = options().spmv_walk;
  // Deflation of the constant vector (generic path): when every row of A sums
  // to the same s > 0 and s is tiny against the diagonal (a graph Laplacian
  // plus a small shift), A 1 = s 1 is a near-null mode that the FP32
  // residual's rounding feeds and an IC0 preconditioner amplifies by ~1/s;
  // the inner iteration is then kept orthogonal to 1 and x's constant
  // component set directly to mean(b) / s.
  bool defl = false;
  double defl_target = 0.0;
  const unsigned gVd = std::min<unsigned>(gV, kMaxPartialBlocks / 3);
  if (!jac && n > 0) {
    constexpr unsigned gs = 256;
    double *st = scratch_vec("mx_rowstats", 4 * gs);
    k_mx_rowstats<kBlock><<<gs, kBlock, 0, stream()>>>(n, ia, ja, a, b, st);
    std::vector<double> h(4 * gs);
    if (LAP_HIP(hipMemcpy(h.data(), st, sizeof(double) * h.size(), hipMemcpyDeviceToHost))) {
      double lo = INFINITY, hi = -INFINITY, dmax = 0.0, sb = 0.0;
      for (unsigned k = 0; k < gs; ++k) {
        lo = std::min(lo, h[4 * k]);
        hi = std::max(hi, h[4 * k + 1]);
        dmax = std::max(dmax, h[4 * k + 2]);
        sb += h[4 * k + 3];
      }
      if (lo > 0.0 && hi - lo <= 1e-6 * hi && hi <= 1e-6 * dmax) {
        defl = true;
        defl_target = (sb / n) / (0.5 * (lo + hi));
      }
    }
  }

  MixCtl h0{};
  h0.maxit = maxit;
  LAP_HIP(hipMemcpyAsync(ctl, &h0, sizeof h0, hipMemcpyHostToDevice, stream()));
  hipEvent_t ev0, ev1;
  LAP_HIP(hipEventCreate(&ev0));
  LAP_HIP(hipEventCreate(&ev1));
  LAP_HIP(hipEventRecord(ev0, stream()));
  k_zero_v<VT><<<gV, kBlock, 0, stream()>>>(n, x32);
  if (defl) {  // x's constant component is known: set it
    k_mx_xsum<kBlock><<<gV, kBlock, 0, stream()>>>(n, x, Q3);
    k_mx_xshift<kBlock><<<gV, kBlock, 0, stream()>>>(n, x, defl_target, ctl, nullptr, 0, Q3,
                                                     (int)gV);
  }
  k_mx_init<kBlock, kSpmvTile, VT><<<gA, kBlock, 0, stream()>>>(
      A64->nblocks, A64->rowblk->as<int>(), ia, ja, a, x, b, r32, d_r32, jac ? p32 : nullptr,
      Q0, walk);
  k_mx_start<kBlock><<<1, kBlock, 0, stream()>>>(Q0, (int)gA, tol, delta, ctl, hist);

  const int *rb_o = M->off->rowblk->as<int>(), *rb_a = A64->rowblk->as<int>();
  const int *io = M->ia_o->as<int>(), *jo = M->ja_o->as<int>();
  const float *ao = M->a32_o->as<float>();
  const double *dg = M->diag->as<double>();
  // value-indexed inner SpMV when spmv_format asks for it and the table fits
  const bool dvi32 = options().spmv_format == 2 && M->ndl > 0;
  const bool vi32 = !dvi32 && options().spmv_format >= 1 && M->nd > 0;
  auto spmv32 = [&]() {
    if (dvi32)
      k_mx_spmv_pq_dvi<kBlock, VT><<<gO, kBlock, 0, stream()>>>(
          M->off->nblocks, rb_o, io, M->code_o->as<unsigned short>(), M->dict_o->as<float>(),
          M->nd, M->ddict_o->as<int>(), M->ndl, dg, p32, q32, ctl, Q1, walk);
    else if (vi32)
      k_mx_spmv_pq_vi<kBlock, kSpmvTile, VT><<<gO, kBlock, 0, stream()>>>(
          M->off->nblocks, rb_o, io, jo, M->vi_o->as<unsigned char>(), M->dict_o->as<float>(),
          M->nd, dg, p32, q32, ctl, Q1, walk);
    else
      k_mx_spmv_pq<kBlock, kSpmvTile, VT><<<gO, kBlock, 0, stream()>>>(M->off->nblocks, rb_o, io,
                                                                   jo, ao, dg, p32, q32, ctl,
                                                                   Q1, walk);
  };
  auto iteration = [&]() {
    if (jac) {
      spmv32();
      k_mx_update<kBlock, true, VT><<<gV, kBlock, 0, stream()>>>(n, x32, r32, p32, q32, d_r32,
                                                                 ctl, Q1, (int)gO, Q2);
      k_mx_resid<kBlock, kSpmvTile, VT><<<gA, kBlock, 0, stream()>>>(
          A64->nblocks, rb_a, ia, ja, a, x, x32, b, r32, d_r32, ctl, Q2, (int)gV, Q3, walk);
      k_mx_jac_p<kBlock, VT><<<gV, kBlock, 0, stream()>>>(n, r32, d_r32, p32, x, x32, ctl, hist,
                                                          Q2, (int)gV, Q3, (int)gA);
    } else {
      if (ic) {
        ic_inner(*icp, *M, l32, u32, r32, w32, aux32, done);
      } else {
        k_copy_v<VT><<<gV, kBlock, 0, stream()>>>(n, r32, w32, done);
      }
      if (defl) {
        k_mx_rw<kBlock, VT, true><<<gVd, kBlock, 0, stream()>>>(n, r32, w32, ctl, Q0);
        k_mx_pupd<kBlock, VT, true><<<gVd, kBlock, 0, stream()>>>(n, w32, p32, ctl, Q0,
                                                                  (int)gVd);
      } else {
        k_mx_rw<kBlock, VT, false><<<gV, kBlock, 0, stream()>>>(n, r32, w32, ctl, Q0);
        k_mx_pupd<kBlock, VT, false><<<gV, kBlock, 0, stream()>>>(n, w32, p32, ctl, Q0,
                                                                  (int)gV);
      }
      spmv32();
      k_mx_update<kBlock, false, VT><<<gV, kBlock, 0, stream()>>>(n, x32, r32, p32, q32,
                                                                  nullptr, ctl, Q1, (int)gO, Q2);
      k_mx_fold<kBlock, VT><<<gV, kBlock, 0, stream()>>>(n, x, x32, ctl, Q2, (int)gV,
                                                         defl ? Q3 : nullptr);
      if (defl)
        k_mx_xshift<kBlock><<<gV, kBlock, 0, stream()>>>(n, x, defl_target, ctl, Q2, (int)gV, Q3,
                                                         (int)gV);
      k_mx_resid<kBlock, kSpmvTile, VT><<<gA, kBlock, 0, stream()>>>(
          A64->nblocks, rb_a, ia, ja, a, x, nullptr, b, r32, nullptr, ctl, Q2, (int)gV, Q3,
          walk);
      k_mx_conv<kBlock><<<1, kBlock, 0, stream()>>>(ctl, hist, Q2, (int)gV, Q3, (int)gA);
    }
  };
  int per_iter = jac ? 4 : defl ? 9 : 8;
  if (ic)
    per_iter += sizeof(VT) == sizeof(double)
                    ? tri_launches(*icp->lplan, &icp->lell, false) +
                          tri_launches(*icp->uplan, &icp->uell, false)
                    : tri_launches32(*icp->lplan, &M->lch) + tri_launches32(*icp->uplan, &M->uch);
  int G = std::max(1, std::min(options().graph_iters, 4096 / per_iter));
  G = std::min(G, maxit + 2);

  int status = LAP_OK;
  hipGraphExec_t gexec = nullptr;
  if (options().use_graph) {
    hipGraph_t graph = nullptr;
    bool ok = LAP_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
    if (ok) {
      for (int k = 0; k < G; ++k) iteration();
      ok = LAP_HIP(hipStreamEndCapture(stream(), &graph));
    }
    if (ok) ok = LAP_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    if (graph) (void)hipGraphDestroy(graph);
    if (!ok) {
      gexec = nullptr;
      status = LAP_ERR_HIP;
    }
  }
  int *flags = pinned_flags();
  if (!flags) status = LAP_ERR_HIP;
  if (status == LAP_OK) {
    hipEvent_t pe[2];
    LAP_HIP(hipEventCreateWithFlags(&pe[0], hipEventDisableTiming));
    LAP_HIP(hipEventCreateWithFlags(&pe[1], hipEventDisableTiming));
    const long max_batches = (long)(maxit + 1) / G + 3;
    bool finished = false;
    for (long k = 0; k < max_batches + 1 && !finished; ++k) {
      if (k < max_batches) {
        if (gexec) {
          if (!LAP_HIP(hipGraphLaunch(gexec, stream()))) break;
        } else {
          for (int s = 0; s < G; ++s) iteration();
        }
        LAP_HIP(hipMemcpyAsync(flags + (k & 1), done, sizeof(int), hipMemcpyDeviceToHost,
                               stream()));
        LAP_HIP(hipEventRecord(pe[k & 1], stream()));
      }
      if (k > 0) {
        LAP_HIP(hipEventSynchronize(pe[(k - 1) & 1]));
        if (flags[(k - 1) & 1]) finished = true;
      }
    }
    if (!finished) {
      set_error(LAP_ERR_NUMERIC, "pcg_mixed: convergence flag never set");
      status = LAP_ERR_NUMERIC;
    }
    (void)hipEventDestroy(pe[0]);
    (void)hipEventDestroy(pe[1]);
  }
  LAP_HIP(hipEventRecord(ev1, stream()));
  sync();
  float ms = 0.f;
  LAP_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  if (gexec) (void)hipGraphExecDestroy(gexec);
  MixCtl h{};
  LAP_HIP(hipMemcpy(&h, ctl, sizeof h, hipMemcpyDeviceToHost));
  out.iterations = h.last_iter;
  out.flag = h.flag;
  out.refinements = h.refinements;
  out.res0 = h.res0;
  out.res = h.res;
  out.gpu_ms = ms;
  if (h_hist)
    LAP_HIP(hipMemcpy(h_hist, hist, sizeof(double) * (size_t)(h.last_iter + 1),
                      hipMemcpyDeviceToHost));
  if (status == LAP_OK && lap_last_status() == LAP_ERR_HIP) status = LAP_ERR_HIP;
  return status;
}

}  // namespace lap

using namespace lap;

extern "C" int lap_pcg_mixed(int n, int nnz, const int *ia, const int *ja, const double *a,
                             double *x, const double *b, double tol, lap_pdata *prec_data,
                             int maxit, double delta, double *history, lap_solve_info *info) {
  MixOut out;
  if (delta == 0.0) delta = 0.01;
  const int st =
      options().mixed_vectors == 1
          ? run_mixed<double>(n, nnz, ia, ja, a, x, b, tol, prec_data, maxit, delta, history, out)
          : run_mixed<float>(n, nnz, ia, ja, a, x, b, tol, prec_data, maxit, delta, history, out);
  if (info) {
    info->iterations = out.iterations;
    info->flag = st == LAP_OK ? out.flag : 2;
    info->res0 = out.res0;
    info->res_final = out.res;
    info->true_res = out.res;  // the final residual is an FP64 one
    info->gpu_ms = out.gpu_ms;
    info->refinements = out.refinements;
  }
  return st;
}
