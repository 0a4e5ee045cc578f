// pcg.hip -- device-resident preconditioned CG: cg(), lap_pcg(), the
// preconditioners and prec_function().
//
// The recurrence is the reference's (cg.c:5-140): r = -(A x0 - b),
// tolrel = tol*||r0||, then per iteration w = M^{-1} r, rho = r.w,
// p = w (first) or p = beta*p + w, q = A p, alpha = rho/(p.q), x += alpha p,
// r -= alpha q, stop when ||r|| < tolrel (flag 0) or iter > maxit (flag 1).
// Each vector update keeps the reference's rounding (no contraction,
// operands associated as in cg.c/simple_blas.c); only the dot products are
// summed in a different (fixed, tree) order.
//
// Device residency: alpha, beta, rho, the residual history and the
// convergence flag live in a device control block.  Every kernel of an
// iteration first reads ctl->done and returns if it is set, so the host can
// launch a HIP graph holding G iterations and poll the flag once per graph:
// the iteration count stays exact.
//
// Jacobi (line_jacobi) iterations are three kernels:
//   K_A  q = A p with p.q fused into the SpMV epilogue; alpha by the last block
//   K_B  x += alpha p, r -= alpha q, w = d_r r, sums r.w and r.r; the last
//        block does the convergence test and beta
//   K_C  p = beta p + d_r r
// Other preconditioners run their own launches followed by r.w, the p
// update, K_A and the x/r update.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "lap_core.hpp"
#include "spmv_core.hpp"

namespace lap {

struct PcgCtl {
  double rho;       // r_k.w_k of the iteration in flight (alpha = rho / p.q)
  double rho_prev;  // generic path: rho of the previous iteration
  double tolrel, res0, res, pend_rho;
  double alpha_x;   // Jacobi: alpha of the x update still to apply
  int iter;         // finished iterations, as seen by the kernels
  int maxit, done, flag;
  int pend_iter, pend_valid, last_iter, pend_done;
};

enum PrecKind { P_NONE, P_JACOBI, P_IT_JACOBI, P_GS_STD, P_GS_IT, P_GS_IT2, P_ICHOL, P_UNKNOWN };

static PrecKind prec_kind(const char *s) {
  if (!s) return P_UNKNOWN;
  if (!strcmp(s, "none")) return P_NONE;
  if (!strcmp(s, "line_jacobi")) return P_JACOBI;
  if (!strcmp(s, "it_jacobi")) return P_IT_JACOBI;
  if (!strcmp(s, "GS_std")) return P_GS_STD;
  if (!strcmp(s, "GS_it")) return P_GS_IT;
  if (!strcmp(s, "GS_it2")) return P_GS_IT2;
  if (!strcmp(s, "ichol")) return P_ICHOL;
  return P_UNKNOWN;
}

namespace {

// partial-sum regions of the reduction workspace (kMaxPartialBlocks each)
constexpr int kR0 = 0, kR1 = kMaxPartialBlocks, kR2 = 2 * kMaxPartialBlocks,
              kR3 = 3 * kMaxPartialBlocks;

inline unsigned vgrid(int n) {
  return grid_for((size_t)(n > 0 ? n : 1), 4, kBlock, (unsigned)options().vec_grid);
}

// ---------------------------------------------------------------- kernels
//
// Control-block discipline: inside one launch a field of PcgCtl is either
// read by every workgroup or written by workgroup 0, never both (except
// `done`, which workgroup 0 only sets when every workgroup reaches the same
// stop decision from the same bits).  Values needed by all workgroups come
// from per-workgroup partials of the previous launch, summed redundantly.

struct InitEpi {
  double *r;
  const double *rhs;
  double acc;
  __device__ double init(int) const { return 0.0; }
  __device__ void store(int row, double s) {
    const double v = -(s - rhs[row]);  // r = A x; r += -1*b; r *= -1  (cg.c:48-55)
    r[row] = v;
    acc += v * v;
  }
};

// r = b - A x0; partials of r.r
template <int BS, int TILE>
__global__ __launch_bounds__(BS) void k_init(int nblk, const int *__restrict__ rowblk,
                                             const int *__restrict__ ia,
                                             const int *__restrict__ ja,
                                             const double *__restrict__ a,
                                             const double *__restrict__ x,
                                             const double *__restrict__ rhs,
                                             double *__restrict__ r, double *partials, int xcd) {
  __shared__ double prod[TILE];
  __shared__ double red[BS / kWave];
  InitEpi epi{r, rhs, 0.0};
  spmv_walk<BS, TILE>(block_walk(nblk, xcd), rowblk, ia, ja, a, x, 1.0, prod, red, epi);
  const double t = block_sum<BS>(epi.acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// ||r0||, tolrel, hist[0].  Jacobi: also w = d_r r, p = w, partials of r.w.
template <int BS, bool JAC>
__global__ __launch_bounds__(BS) void k_start(int n, const double *rr_part, int rr_cnt,
                                              double tol, const double *__restrict__ r,
                                              const double *__restrict__ d_r,
                                              double *__restrict__ p, PcgCtl *ctl,
                                              double *hist, double *rz_part) {
  __shared__ double red[BS / kWave];
  const double res0 = sqrt(sum_partials<BS>(rr_part, rr_cnt, 1, red));
  const bool stop = !isfinite(res0) || res0 == 0.0;  // res0 == 0: x0 solves the system
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    hist[0] = res0;
    ctl->res0 = res0;
    ctl->res = res0;
    ctl->tolrel = tol * res0;
    if (stop) {
      ctl->flag = isfinite(res0) ? 0 : 2;
      ctl->done = 1;
      ctl->pend_done = 1;
    }
  }
  if (!JAC || stop) return;
  double s = 0.0;
  const size_t stride = (size_t)gridDim.x * BS;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride) {
    const double ri = r[i];
    const double wi = d_r[i] * ri;  // line_jacobi: w = d_r .* r
    p[i] = wi;                      // first iteration: p = w
    s += ri * wi;
  }
  const double t = block_sum<BS>(s, red);
  if (threadIdx.x == 0) rz_part[blockIdx.x] = t;
}

// one workgroup: rho = sum of the r.w partials
template <int BS>
__global__ __launch_bounds__(BS) void k_rho_commit(const double *part, int cnt, PcgCtl *ctl) {
  if (ctl->done) return;
  __shared__ double red[BS / kWave];
  const double rho = sum_partials<BS>(part, cnt, 1, red);
  if (threadIdx.x == 0) {
    ctl->rho = rho;
    ctl->pend_rho = rho;
  }
}

struct PqEpi {
  double *q;
  const double *p;
  double acc;
  __device__ double init(int) const { return 0.0; }
  __device__ void store(int row, double s) {
    st_nt(q + row, s);  // q is read once, by the next kernel
    acc += p[row] * s;
  }
};

// K_A: q = A p, partials of p.q.  Workgroup 0 commits the pending
// iteration count and rho left by the previous iteration's last kernel.
template <int BS, int TILE>
__global__ __launch_bounds__(BS) void k_spmv_pq(int nblk, const int *__restrict__ rowblk,
                                                const int *__restrict__ ia,
                                                const int *__restrict__ ja,
                                                const double *__restrict__ a,
                                                const double *__restrict__ p,
                                                double *__restrict__ q, PcgCtl *ctl,
                                                double *partials, int xcd) {
  if (ctl->done) return;
  if (ctl->pend_done) {  // the previous kernel stopped the solve: commit it
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->done = 1;
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && ctl->pend_valid) {
    ctl->iter = ctl->pend_iter;
    ctl->rho = ctl->pend_rho;
    ctl->pend_valid = 0;
  }
  __shared__ double prod[TILE];
  __shared__ double red[BS / kWave];
  PqEpi epi{q, p, 0.0};
  spmv_walk<BS, TILE>(block_walk(nblk, xcd), rowblk, ia, ja, a, p, 1.0, prod, red, epi);
  const double t = block_sum<BS>(epi.acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// K_A over value-indexed CSR: the same kernel with the values read as
// dict[vi[e]], the table staged in LDS.
template <int BS, int TILE>
__global__ __launch_bounds__(BS) void k_spmv_pq_vi(int nblk, const int *__restrict__ rowblk,
                                                   const int *__restrict__ ia,
                                                   const int *__restrict__ ja,
                                                   const unsigned char *__restrict__ vi,
                                                   const double *__restrict__ dict, int nd,
                                                   const double *__restrict__ p,
                                                   double *__restrict__ q, PcgCtl *ctl,
                                                   double *partials, int xcd) {
  if (ctl->done) return;
  if (ctl->pend_done) {  // the previous kernel stopped the solve: commit it
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->done = 1;
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && ctl->pend_valid) {
    ctl->iter = ctl->pend_iter;
    ctl->rho = ctl->pend_rho;
    ctl->pend_valid = 0;
  }
  __shared__ double prod[TILE];
  __shared__ double red[BS / kWave];
  __shared__ double table[256];
  for (int i = threadIdx.x; i < nd; i += BS) table[i] = dict[i];
  __syncthreads();
  PqEpi epi{q, p, 0.0};
  const BlockWalk w = block_walk(nblk, xcd);
  for (int s = w.begin; s < w.end; s += w.step) {
    const int b = w.block(s);
    if (b >= nblk) continue;
    spmv_block_rows_vi<BS, TILE>(b, rowblk, ia, ja, vi, table, p, 1.0, prod, red, epi);
  }
  const double t = block_sum<BS>(epi.acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

constexpr int kDviTile = kSpmvTile;  // codes staged per row block

// K_A over delta-indexed CSR: nonzero e of row r holds the 16-bit code
// c = v | d << 8, value vt[v] and column r + dt[d] (both tables in LDS).  A
// row block's row pointers and codes are staged into LDS with coalesced
// loads; then one lane per row (rows of at most 64 entries) decodes its
// entries eight at a time and sums the products (1.0 * value) * p[col] in
// CSR order, so each row is bitwise the CSR kernel's.  2 bytes per nonzero
// instead of 12.
template <int BS>
__global__ __launch_bounds__(BS) void k_spmv_pq_dvi(int nblk, const int *__restrict__ rowblk,
                                                    const int *__restrict__ ia,
                                                    const unsigned short *__restrict__ code,
                                                    const double *__restrict__ vdict, int nv,
                                                    const int *__restrict__ ddict, int ndl,
                                                    const double *__restrict__ p,
                                                    double *__restrict__ q, PcgCtl *ctl,
                                                    double *partials, int xcd) {
  if (ctl->done) return;
  if (ctl->pend_done) {  // the previous kernel stopped the solve: commit it
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->done = 1;
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && ctl->pend_valid) {
    ctl->iter = ctl->pend_iter;
    ctl->rho = ctl->pend_rho;
    ctl->pend_valid = 0;
  }
  __shared__ double red[BS / kWave];
  __shared__ double vt[256];
  __shared__ int dt[256];
  __shared__ unsigned short sc[kDviTile];  // the block's codes
  __shared__ int rp[2 * BS + 1];           // its row pointers, relative to e0
  for (int i = threadIdx.x; i < nv; i += BS) vt[i] = vdict[i];
  for (int i = threadIdx.x; i < ndl; i += BS) dt[i] = ddict[i];
  PqEpi epi{q, p, 0.0};
  const BlockWalk w = block_walk(nblk, xcd);
  for (int s = w.begin; s < w.end; s += w.step) {
    const int b = w.block(s);
    if (b >= nblk) continue;
    const int r0 = rowblk[b], r1 = rowblk[b + 1];
    const int e0 = ia[r0];
    // coalesced staging: row pointers, then the codes of the block
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
        for (int j = 0; j < 8; ++j) xv[j] = k + j < k1 ? p[row + dt[c[j] >> 8]] : 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (k + j < k1) sum += (1.0 * vt[c[j] & 255u]) * xv[j];
      }
      epi.store(row, sum);
    }
    __syncthreads();
  }
  const double t = block_sum<BS>(epi.acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// A value-indexed copy of the matrix for one solve (vi == nullptr and
// code == nullptr: the CSR arrays are used).
struct ViVals {
  const unsigned char *vi = nullptr;
  const double *dict = nullptr;
  int nd = 0;
  const unsigned short *code = nullptr;  // delta-indexed (dict = value table)
  const int *ddict = nullptr;
  int ndl = 0;
};

constexpr int kDviMaxRow = 64;

// Encodes the matrix when options().spmv_format asks for it and it fits:
// format 2 tries the delta-indexed form first, then (like format 1) the
// value-indexed one.
static ViVals vi_prepare(int nnz, const double *a, int n = 0, const int *ia = nullptr,
                         const int *ja = nullptr) {
  ViVals v;
  const int fmt = options().spmv_format;
  if (fmt == 2 && ia && ja && n > 0) {
    unsigned short *code =
        reinterpret_cast<unsigned short *>(scratch_vec("dvi_code", ((size_t)nnz + 3) / 4 + 1));
    double *vd = scratch_vec("dvi_vdict", 256);
    int *dd = reinterpret_cast<int *>(scratch_vec("dvi_ddict", 128));
    int nv = 0, ndl = 0;
    if (dvi_encode(n, nnz, ia, ja, a, kDviMaxRow, code, vd, dd, &nv, &ndl) == 1) {
      v.code = code;
      v.dict = vd;
      v.nd = nv;
      v.ddict = dd;
      v.ndl = ndl;
      return v;
    }
  }
  if (fmt != 1 && fmt != 2) return v;
  unsigned char *vi =
      reinterpret_cast<unsigned char *>(scratch_vec("vi_idx", ((size_t)nnz + 7) / 8 + 1));
  double *dict = scratch_vec("vi_dict", 256);
  const int nd = vi_encode(nnz, a, vi, dict);
  if (nd > 0) {
    v.vi = vi;
    v.dict = dict;
    v.nd = nd;
  }
  return v;
}

// Convergence decision of iteration `it` (cg.c:106-123): -1 keep going,
// 0 converged, 1 maxit, 2 non-finite residual.
__device__ __forceinline__ int stop_flag(double res, int it, const PcgCtl *ctl) {
  if (!isfinite(res)) return 2;
  if (res < ctl->tolrel) return 0;
  if (it > ctl->maxit) return 1;
  return -1;
}

// ------------------------------------------------------ Jacobi iteration
//
// The x update x += alpha p is done in K_C, which reads p anyway for
// p = beta p + d_r r (p is read once per iteration less than with x updated
// beside r).  K_B leaves alpha in ctl->alpha_x; when K_C finds that the
// iteration stops it applies only the x update, and its workgroup 0 marks
// pend_done (committed to done by the next K_A), so that no workgroup of K_C
// can skip its rows' x update by seeing `done` early.  Same operations on the
// same values as cg.c's order of updates.

// K_B: alpha = rho/(p.q); r -= alpha q; partials of r.(d_r r) and r.r
template <int BS>
__global__ __launch_bounds__(BS) void k_jac_update_nx(int n, double *__restrict__ r,
                                                      const double *__restrict__ q,
                                                      const double *__restrict__ d_r,
                                                      PcgCtl *ctl, const double *pq_part,
                                                      int pq_cnt, double *partials) {
  if (ctl->done) return;
  __shared__ double red[BS / kWave];
  const double pq = sum_partials<BS>(pq_part, pq_cnt, 1, red);
  const double alpha = ctl->rho / pq, nalpha = -1.0 * alpha;
  double rz = 0.0, rr = 0.0;
  const size_t stride = (size_t)gridDim.x * BS;
  // every vector here is streamed once (non-temporal: 256^3 Jacobi PCG 575
  // against 610 us per iteration with ordinary loads and stores)
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride) {
    const double ri = ld_nt(r + i) + nalpha * ld_nt(q + i);
    st_nt(r + i, ri);
    const double wi = ld_nt(d_r + i) * ri;
    rz += ri * wi;
    rr += ri * ri;
  }
  const double t_rz = block_sum<BS>(rz, red);
  const double t_rr = block_sum<BS>(rr, red);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = t_rz;
    partials[2 * blockIdx.x + 1] = t_rr;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->alpha_x = alpha;
}

// K_C: convergence test and beta; x += alpha p, then p = beta p + d_r r
template <int BS>
__global__ __launch_bounds__(BS) void k_jac_px(int n, double *__restrict__ x,
                                               const double *__restrict__ r,
                                               const double *__restrict__ d_r,
                                               double *__restrict__ p, PcgCtl *ctl,
                                               double *hist, const double *part, int cnt) {
  if (ctl->done) return;
  __shared__ double red[BS / kWave];
  const double rz = sum_partials<BS>(part, cnt, 2, red);
  const double rr = sum_partials<BS>(part + 1, cnt, 2, red);
  const int it = ctl->iter + 1;
  const double res = sqrt(rr);
  const int flag = stop_flag(res, it, ctl);
  const double alpha = ctl->alpha_x;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    hist[it] = res;
    ctl->res = res;
    ctl->last_iter = it;
    if (flag >= 0) {
      ctl->flag = flag;
      ctl->pend_done = 1;
    } else {
      ctl->pend_iter = it;
      ctl->pend_rho = rz;
      ctl->pend_valid = 1;
    }
  }
  const size_t stride = (size_t)gridDim.x * BS;
  if (flag >= 0) {  // the last iteration: only its x update remains
    for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride)
      x[i] += alpha * p[i];
    return;
  }
  const double beta = rz / ctl->rho;
  // x += alpha*p (cg.c:87); p *= beta; p += 1.0*w (cg.c:80-81)
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride) {
    const double pi = ld_nt(p + i);
    st_nt(x + i, ld_nt(x + i) + alpha * pi);
    st_nt(p + i, pi * beta + ld_nt(d_r + i) * ld_nt(r + i));
  }
}

// generic: partials of r.w
template <int BS>
__global__ __launch_bounds__(BS) void k_rw(int n, const double *__restrict__ r,
                                           const double *__restrict__ w, const PcgCtl *ctl,
                                           double *partials) {
  if (ctl->done) return;
  __shared__ double red[BS / kWave];
  double s = 0.0;
  const size_t stride = (size_t)gridDim.x * BS;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride)
    s += ld_nt(r + i) * ld_nt(w + i);
  const double t = block_sum<BS>(s, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// generic: rho = r.w; p = w (first iteration) or p = beta p + w
template <int BS>
__global__ __launch_bounds__(BS) void k_pupd(int n, const double *__restrict__ w,
                                             double *__restrict__ p, PcgCtl *ctl,
                                             const double *part, int cnt) {
  if (ctl->done) return;
  __shared__ double red[BS / kWave];
  const double rho = sum_partials<BS>(part, cnt, 1, red);
  const bool first = ctl->iter == 0;
  const double beta = first ? 0.0 : rho / ctl->rho_prev;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->rho = rho;
  const size_t stride = (size_t)gridDim.x * BS;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride)
    st_nt(p + i, first ? ld_nt(w + i) : ld_nt(p + i) * beta + ld_nt(w + i));
}

// generic: alpha = rho/(p.q); x += alpha p; r -= alpha q; partials of r.r
template <int BS>
__global__ __launch_bounds__(BS) void k_update(int n, double *__restrict__ x,
                                               double *__restrict__ r,
                                               const double *__restrict__ p,
                                               const double *__restrict__ q,
                                               const PcgCtl *ctl, const double *pq_part,
                                               int pq_cnt, double *partials) {
  if (ctl->done) return;
  __shared__ double red[BS / kWave];
  const double pq = sum_partials<BS>(pq_part, pq_cnt, 1, red);
  const double alpha = ctl->rho / pq, nalpha = -1.0 * alpha;
  double rr = 0.0;
  const size_t stride = (size_t)gridDim.x * BS;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride) {
    st_nt(x + i, ld_nt(x + i) + alpha * ld_nt(p + i));
    const double ri = ld_nt(r + i) + nalpha * ld_nt(q + i);
    st_nt(r + i, ri);
    rr += ri * ri;
  }
  const double t = block_sum<BS>(rr, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// generic, one workgroup: ||r||, convergence test, bookkeeping
template <int BS>
__global__ __launch_bounds__(BS) void k_conv(PcgCtl *ctl, double *hist, const double *part,
                                             int cnt) {
  if (ctl->done) return;
  __shared__ double red[BS / kWave];
  const double rr = sum_partials<BS>(part, cnt, 1, red);
  if (threadIdx.x == 0) {
    const int it = ctl->iter + 1;
    const double res = sqrt(rr);
    const int flag = stop_flag(res, it, ctl);
    hist[it] = res;
    ctl->res = res;
    ctl->iter = it;
    ctl->last_iter = it;
    ctl->rho_prev = ctl->rho;
    if (flag >= 0) {
      ctl->flag = flag;
      ctl->done = 1;
    }
  }
}

}  // namespace

// SpMV-based PCG kernels
static void launch_init(unsigned g, const SpmvPlan &A, const int *ia, const int *ja,
                        const double *a, const double *x, const double *b, double *r,
                        double *part) {
  const int xcd = options().spmv_walk;
  const int *rb = A.rowblk->as<int>();
  k_init<kBlock, kSpmvTile><<<g, kBlock, 0, stream()>>>(A.nblocks, rb, ia, ja, a, x, b, r, part,
                                                     xcd);
}

static void launch_spmv_pq(unsigned g, const SpmvPlan &A, const int *ia, const int *ja,
                           const double *a, const double *p, double *q, PcgCtl *ctl,
                           double *part, const ViVals &vv = ViVals()) {
  const int xcd = options().spmv_walk;
  const int *rb = A.rowblk->as<int>();
  if (vv.code) {
    k_spmv_pq_dvi<kBlock><<<g, kBlock, 0, stream()>>>(A.nblocks, rb, ia, vv.code, vv.dict, vv.nd,
                                                      vv.ddict, vv.ndl, p, q, ctl, part, xcd);
    return;
  }
  if (vv.vi) {
    k_spmv_pq_vi<kBlock, kSpmvTile><<<g, kBlock, 0, stream()>>>(A.nblocks, rb, ia, ja, vv.vi,
                                                                vv.dict, vv.nd, p, q, ctl, part,
                                                                xcd);
    return;
  }
  k_spmv_pq<kBlock, kSpmvTile><<<g, kBlock, 0, stream()>>>(A.nblocks, rb, ia, ja, a, p, q, ctl,
                                                        part, xcd);
}

// ------------------------------------------------------ preconditioner ops

struct PrecOps {
  PrecKind kind = P_UNKNOWN;
  int n = 0, nnzA = 0, m = 1, k = 1;
  const int *ia = nullptr, *ja = nullptr;
  const double *a = nullptr;
  lap_pdata *pd = nullptr;
  const SpmvPlan *A = nullptr, *Ls = nullptr, *Us = nullptr;  // SpMV plans
  const PairPlan *LU = nullptr;                               // L and U together
  const TriPlan *Lt = nullptr, *Ut = nullptr;                 // triangular plans
  const double *ldiag = nullptr, *udiag = nullptr;
  TriEll lell, uell;  // GS_std: the triangles packed for this solve (run_pcg)
  const IcPlan *ic = nullptr;
  double *aux1 = nullptr, *aux2 = nullptr, *aux3 = nullptr;
};

static bool check_dev(const void *p, const char *what) {
  if (is_device_ptr(p)) return true;
  set_error(LAP_ERR_ARG, "%s is not device memory (%p)", what, p);
  return false;
}

// Resolves plans and validates every operand the preconditioner will read.
static bool prec_setup(PrecOps &o, int n, int nnzA, const int *ia, const int *ja,
                       const double *a, lap_pdata *pd) {
  o.kind = prec_kind(pd ? pd->prec_op : nullptr);
  o.n = n;
  o.nnzA = nnzA;
  o.ia = ia;
  o.ja = ja;
  o.a = a;
  o.pd = pd;
  if (o.kind == P_UNKNOWN) {
    set_error(LAP_ERR_PREC, "unknown preconditioner '%s'", pd && pd->prec_op ? pd->prec_op : "(null)");
    return false;
  }
  if (pd->m < 0 || pd->k < 0) {
    set_error(LAP_ERR_ARG, "negative M/K (%d, %d)", pd->m, pd->k);
    return false;
  }
  o.m = pd->m;
  o.k = pd->k;
  o.A = spmv_plan(n, nnzA, ia, ja);
  if (!o.A) return false;
  auto aux = [&](double *given, const char *tag) -> double * {
    if (given && is_device_ptr(given)) {
      hipDeviceptr_t base = nullptr;
      size_t size = 0;
      if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)given) == hipSuccess &&
          (size_t)((const char *)base + size - (const char *)given) >= sizeof(double) * (size_t)n)
        return given;
      (void)hipGetLastError();
    }
    return scratch_vec(tag, (size_t)n);
  };
  o.aux1 = aux(pd->aux_vec1, "prec_aux1");
  o.aux2 = aux(pd->aux_vec2, "prec_aux2");
  o.aux3 = aux(pd->aux_vec3, "prec_aux3");
  const bool need_dr = o.kind == P_JACOBI || o.kind == P_IT_JACOBI || o.kind == P_GS_IT ||
                       o.kind == P_GS_IT2;
  if (need_dr && !check_device("preconditioner",
                               {{pd->d_r, "prec_data->d_r", sizeof(double) * (size_t)n}}))
    return false;
  if (o.kind == P_IT_JACOBI || o.kind == P_GS_IT || o.kind == P_GS_IT2) {
    if (!check_device("preconditioner", {{pd->lia, "prec_data->lia"}, {pd->lja, "prec_data->lja"},
                                         {pd->la, "prec_data->la"}, {pd->uia, "prec_data->uia"},
                                         {pd->uja, "prec_data->uja"}, {pd->ua, "prec_data->ua"}}))
      return false;
    o.Ls = spmv_plan(n, pd->lnnz, pd->lia, pd->lja);
    o.Us = spmv_plan(n, pd->unnz, pd->uia, pd->uja);
    if (!o.Ls || !o.Us) return false;
    if (o.kind != P_GS_IT) o.LU = pair_plan(n, pd->lnnz, pd->lia, pd->unnz, pd->uia);
  }
  if (o.kind == P_GS_STD) {
    if (!check_device("GS_std", {{pd->lia, "prec_data->lia"}, {pd->lja, "prec_data->lja"},
                                 {pd->la, "prec_data->la"}, {pd->uia, "prec_data->uia"},
                                 {pd->uja, "prec_data->uja"}, {pd->ua, "prec_data->ua"}}))
      return false;
    tri_plan_pair(n, pd->lnnz, pd->lia, pd->lja, pd->unnz, pd->uia, pd->uja);
    o.Lt = tri_plan(n, pd->lnnz, pd->lia, pd->lja, true);
    o.Ut = tri_plan(n, pd->unnz, pd->uia, pd->uja, false);
    if (!o.Lt || !o.Ut) return false;
    // pivots live in the matrices (GPU driver layout) or in prec_data->d
    if (!o.Lt->has_diag || !o.Ut->has_diag) {
      if (!check_dev(pd->d, "prec_data->d (strict L/U need the diagonal)")) return false;
      o.ldiag = o.Lt->has_diag ? nullptr : pd->d;
      o.udiag = o.Ut->has_diag ? nullptr : pd->d;
    }
  }
  if (o.kind == P_ICHOL) {
    o.ic = ic_plan_find(pd->ichol_vals);
    if (!o.ic) {
      set_error(LAP_ERR_PREC, "ichol: no factor for ichol_vals=%p (call initialize_ichol)",
                (void *)pd->ichol_vals);
      return false;
    }
  }
  return true;
}

static void spmv_l(const PrecOps &o, const double *x, double *y, double al, double be,
                   const int *done) {
  spmv_launch(*o.Ls, o.pd->lia, o.pd->lja, o.pd->la, x, y, al, be, done);
}
static void spmv_u(const PrecOps &o, const double *x, double *y, double al, double be,
                   const int *done) {
  spmv_launch(*o.Us, o.pd->uia, o.pd->uja, o.pd->ua, x, y, al, be, done);
}

// The SpMV-only preconditioners with one launch per sweep (sweep.hip).
// Every value is bitwise the unfused sequence in prec_launch for finite
// operands, with two shortcuts that use x = 0: it_jacobi's first sweep is
// d_r .* r, and GS_it's first residual r - A*0 is r itself (the unfused sum
// can turn a -0.0 entry of r into +0.0; no value changes).  Returns false
// (the caller runs the unfused sequence) for M or K below 1 and when w
// shares memory with r or an aux vector.
static bool prec_fused(const PrecOps &o, const double *r, double *w, const int *done) {
  if (!options().fuse_sweeps || o.k < 1 || o.m < 1) return false;
  if (o.kind != P_IT_JACOBI && o.kind != P_GS_IT && o.kind != P_GS_IT2) return false;
  if (w == r || w == o.aux1 || w == o.aux2 || w == o.aux3) return false;
  if (o.kind != P_GS_IT && !o.LU) return false;
  const int n = o.n, K = o.k;
  lap_pdata *pd = o.pd;
  const double *dr = pd->d_r;
  auto pair = [&](const double *x, const double *base, double *out) {
    sweep_pair_launch(*o.LU, pd->lia, pd->lja, pd->la, pd->uia, pd->uja, pd->ua, x, base, dr,
                      out, done);
  };
  auto one_l = [&](const double *x, const double *base, double *out, double *out2, int mode,
                   bool first) {
    sweep_one_launch(*o.Ls, pd->lia, pd->lja, pd->la, x, base, dr, out, out2, mode, first, done);
  };
  auto one_u = [&](const double *x, const double *base, double *out, double *out2, int mode,
                   bool first) {
    sweep_one_launch(*o.Us, pd->uia, pd->uja, pd->ua, x, base, dr, out, out2, mode, first, done);
  };
  if (o.kind == P_IT_JACOBI) {  // it_jacobi.c:4-32
    // sweep s (1..K) writes w when K - s is even, else aux1, so sweep K
    // ends in w
    auto dst = [&](int s) { return ((K - s) % 2 == 0) ? w : o.aux1; };
    launch_vec_vec(n, r, dr, dst(1), done);
    for (int s = 2; s <= K; ++s) pair(dst(s - 1), r, dst(s));
    return true;
  }
  if (o.kind == P_GS_IT) {  // GS.c:45-92
    for (int j = 0; j < o.m; ++j) {
      const double *base = r;  // r - A w (w = 0 at j = 0)
      if (j == 0)
        launch_vec_vec(n, r, dr, o.aux1, done);
      else {
        sweep_one_launch(*o.A, o.ia, o.ja, o.a, w, r, dr, o.aux2, o.aux1, 2, false, done);
        base = o.aux2;
      }
      double *y = o.aux1, *alt = o.aux3;
      for (int s = 0; s < 2 * K; ++s) {
        const bool lower = s < K;
        if (s == 2 * K - 1) {  // last U sweep: w += y .* d_r
          one_u(y, base, nullptr, w, 3, j == 0);
        } else {
          if (lower)
            one_l(y, base, alt, nullptr, 0, false);
          else
            one_u(y, base, alt, nullptr, 0, false);
          std::swap(y, alt);
        }
      }
    }
    return true;
  }
  // GS_it2, GS.c:95-159: y ping-pongs between aux1 and aux2, aux3 holds the
  // residuals r - L y and r - U y; the last sweep writes w
  double *buf[2] = {o.aux1, o.aux2};
  int cur = 0;
  launch_vec_vec(n, r, dr, buf[0], done);
  for (int j = 0; j < o.m; ++j) {
    pair(buf[cur], r, buf[cur ^ 1]);
    cur ^= 1;
    one_l(buf[cur], r, o.aux3, nullptr, 1, false);
    for (int s = 0; s < K; ++s) {
      one_u(buf[cur], o.aux3, buf[cur ^ 1], nullptr, 0, false);
      cur ^= 1;
    }
    one_u(buf[cur], r, o.aux3, nullptr, 1, false);
    for (int s = 0; s < K; ++s) {
      const bool last = j == o.m - 1 && s == K - 1;
      one_l(buf[cur], o.aux3, last ? w : buf[cur ^ 1], nullptr, 0, false);
      cur ^= 1;
    }
  }
  return true;
}

// w = M^{-1} r, as launches on the library stream.
static void prec_launch(const PrecOps &o, const double *r, double *w, const int *done) {
  const int n = o.n;
  lap_pdata *pd = o.pd;
  if (prec_fused(o, r, w, done)) return;
  switch (o.kind) {
    case P_NONE:
      launch_copy(n, r, w, done);
      break;
    case P_JACOBI:  // line_jacobi.c:3-7
      launch_vec_vec(n, pd->d_r, r, w, done);
      break;
    case P_IT_JACOBI:  // it_jacobi.c:4-32
      launch_fill(n, 0.0, w, done);
      for (int s = 0; s < o.k; ++s) {
        spmv_l(o, w, o.aux1, 1.0, 0.0, done);
        spmv_u(o, w, o.aux2, 1.0, 0.0, done);
        launch_axpy(n, 1.0, o.aux1, o.aux2, done);
        launch_copy(n, r, w, done);
        launch_axpy(n, -1.0, o.aux2, w, done);
        launch_vec_vec(n, w, pd->d_r, w, done);
      }
      break;
    case P_GS_STD:  // GS.c:5-42
      if (o.k == 0) launch_fill(n, 0.0, w, done);
      for (int s = 0; s < o.k; ++s) {
        if (s == 0) {
          // w = 0: r - A*0 == r and 0 + L^{-1} r == L^{-1} r, so solve directly
          tri_solve_launch(*o.Lt, pd->lia, pd->lja, pd->la, o.ldiag, r, w, done, &o.lell);
        } else {
          launch_copy(n, r, o.aux2, done);
          spmv_launch(*o.A, o.ia, o.ja, o.a, w, o.aux2, -1.0, 1.0, done);
          tri_solve_launch(*o.Lt, pd->lia, pd->lja, pd->la, o.ldiag, o.aux2, o.aux1, done,
                           &o.lell);
          launch_axpy(n, 1.0, o.aux1, w, done);
        }
      }
      for (int s = 0; s < o.k; ++s) {
        launch_copy(n, r, o.aux2, done);
        spmv_launch(*o.A, o.ia, o.ja, o.a, w, o.aux2, -1.0, 1.0, done);
        tri_solve_launch(*o.Ut, pd->uia, pd->uja, pd->ua, o.udiag, o.aux2, o.aux1, done,
                         &o.uell);
        launch_axpy(n, 1.0, o.aux1, w, done);
      }
      break;
    case P_GS_IT:  // GS.c:45-92
      launch_fill(n, 0.0, w, done);
      for (int j = 0; j < o.m; ++j) {
        launch_copy(n, r, o.aux2, done);
        spmv_launch(*o.A, o.ia, o.ja, o.a, w, o.aux2, -1.0, 1.0, done);
        launch_vec_vec(n, o.aux2, pd->d_r, o.aux1, done);
        for (int s = 0; s < o.k; ++s) {
          launch_copy(n, o.aux2, o.aux3, done);
          spmv_l(o, o.aux1, o.aux3, -1.0, 1.0, done);
          launch_vec_vec(n, o.aux3, pd->d_r, o.aux1, done);
        }
        for (int s = 0; s < o.k; ++s) {
          launch_copy(n, o.aux2, o.aux3, done);
          spmv_u(o, o.aux1, o.aux3, -1.0, 1.0, done);
          launch_vec_vec(n, o.aux3, pd->d_r, o.aux1, done);
        }
        launch_axpy(n, 1.0, o.aux1, w, done);
      }
      break;
    case P_GS_IT2:  // GS.c:95-159
      launch_vec_vec(n, r, pd->d_r, o.aux1, done);
      for (int j = 0; j < o.m; ++j) {
        spmv_l(o, o.aux1, o.aux2, 1.0, 0.0, done);
        spmv_u(o, o.aux1, o.aux3, 1.0, 0.0, done);
        launch_axpy(n, 1.0, o.aux3, o.aux2, done);
        launch_copy(n, r, o.aux3, done);
        launch_axpy(n, -1.0, o.aux2, o.aux3, done);
        launch_vec_vec(n, o.aux3, pd->d_r, o.aux2, done);
        launch_copy(n, r, o.aux3, done);
        spmv_l(o, o.aux2, o.aux3, -1.0, 1.0, done);
        for (int s = 0; s < o.k; ++s) {
          launch_copy(n, o.aux3, o.aux1, done);
          spmv_u(o, o.aux2, o.aux1, -1.0, 1.0, done);
          launch_vec_vec(n, pd->d_r, o.aux1, o.aux2, done);
        }
        launch_copy(n, r, o.aux3, done);
        spmv_u(o, o.aux2, o.aux3, -1.0, 1.0, done);
        for (int s = 0; s < o.k; ++s) {
          launch_copy(n, o.aux3, o.aux1, done);
          spmv_l(o, o.aux2, o.aux1, -1.0, 1.0, done);
          launch_vec_vec(n, pd->d_r, o.aux1, o.aux2, done);
        }
        launch_copy(n, o.aux2, o.aux1, done);
      }
      launch_copy(n, o.aux2, w, done);
      break;
    case P_ICHOL:  // simple_blas.c:184-223 order
      ic_apply_launch(*o.ic, r, w, o.aux1, done);
      break;
    default:
      break;
  }
}

// ------------------------------------------------------------- the solve

struct SolveOut {
  int iterations = 0, flag = 2;
  double res0 = 0, res = 0, gpu_ms = 0;
};

static int run_pcg(int n, int nnz, const int *ia, const int *ja, const double *a, double *x,
                   const double *b, double tol, lap_pdata *pd, int maxit, double *h_hist,
                   SolveOut &out) {
  if (n <= 0 || nnz < 0 || maxit < 0) {
    set_error(LAP_ERR_ARG, "pcg: bad sizes n=%d nnz=%d maxit=%d", n, nnz, maxit);
    return LAP_ERR_ARG;
  }
  if (!check_device("pcg", {{ia, "ia", sizeof(int) * ((size_t)n + 1)},
                           {ja, "ja", sizeof(int) * (size_t)nnz},
                           {a, "a", sizeof(double) * (size_t)nnz},
                           {x, "x", sizeof(double) * (size_t)n},
                           {b, "b", sizeof(double) * (size_t)n}}))
    return LAP_ERR_ARG;
  PrecOps o;
  if (!prec_setup(o, n, nnz, ia, ja, a, pd)) return lap_last_status();
  if (o.kind == P_GS_STD) {  // packed triangles for the level launches of this solve
    tri_ell_build(*o.Lt, pd->lja, pd->la, o.lell);
    tri_ell_build(*o.Ut, pd->uja, pd->ua, o.uell);
  }
  ReduceWs &ws = reduce_ws();
  double *r = scratch_vec("pcg_r", (size_t)n);
  double *w = scratch_vec("pcg_w", (size_t)n);
  double *p = scratch_vec("pcg_p", (size_t)n);
  double *q = scratch_vec("pcg_q", (size_t)n);
  PcgCtl *ctl = reinterpret_cast<PcgCtl *>(scratch_vec("pcg_ctl", sizeof(PcgCtl) / 8 + 1));
  double *hist = scratch_vec("pcg_hist", (size_t)maxit + 2);
  const int *done = &ctl->done;
  const SpmvPlan &A = *o.A;
  const unsigned gA = spmv_grid(A), gV = vgrid(n);

  double *R0 = ws.partials + kR0, *R1 = ws.partials + kR1, *R2 = ws.partials + kR2,
         *R3 = ws.partials + kR3;
  const bool chunked = options().sptrsv_mode != 0 && (o.kind == P_ICHOL || o.kind == P_GS_STD);
  if (chunked) tri_clear_error();
  PcgCtl h0{};
  h0.maxit = maxit;
  LAP_HIP(hipMemcpyAsync(ctl, &h0, sizeof h0, hipMemcpyHostToDevice, stream()));

  hipEvent_t ev0, ev1;
  LAP_HIP(hipEventCreate(&ev0));
  LAP_HIP(hipEventCreate(&ev1));
  LAP_HIP(hipEventRecord(ev0, stream()));
  launch_init(gA, A, ia, ja, a, x, b, r, R0);
  const bool jac = o.kind == P_JACOBI;
  const ViVals vv = vi_prepare(nnz, a, n, ia, ja);
  if (jac) {
    k_start<kBlock, true><<<gV, kBlock, 0, stream()>>>(n, R0, (int)gA, tol, r, pd->d_r, p, ctl,
                                                       hist, R1);
    k_rho_commit<kBlock><<<1, kBlock, 0, stream()>>>(R1, (int)gV, ctl);
  } else {
    k_start<kBlock, false><<<gV, kBlock, 0, stream()>>>(n, R0, (int)gA, tol, r, nullptr, nullptr,
                                                        ctl, hist, nullptr);
  }
  auto iteration = [&]() {
    if (jac) {
      launch_spmv_pq(gA, A, ia, ja, a, p, q, ctl, R2, vv);
      k_jac_update_nx<kBlock><<<gV, kBlock, 0, stream()>>>(n, r, q, pd->d_r, ctl, R2, (int)gA,
                                                           R3);
      k_jac_px<kBlock><<<gV, kBlock, 0, stream()>>>(n, x, r, pd->d_r, p, ctl, hist, R3, (int)gV);
    } else {
      prec_launch(o, r, w, done);
      k_rw<kBlock><<<gV, kBlock, 0, stream()>>>(n, r, w, ctl, R1);
      k_pupd<kBlock><<<gV, kBlock, 0, stream()>>>(n, w, p, ctl, R1, (int)gV);
      launch_spmv_pq(gA, A, ia, ja, a, p, q, ctl, R2, vv);
      k_update<kBlock><<<gV, kBlock, 0, stream()>>>(n, x, r, p, q, ctl, R2, (int)gA, R3);
      k_conv<kBlock><<<1, kBlock, 0, stream()>>>(ctl, hist, R3, (int)gV);
    }
  };
  // launches per iteration -> iterations per graph.  Jacobi stops through
  // pend_done and is polled there (done follows in the next launch).
  const int *poll = jac ? &ctl->pend_done : done;
  int per_iter = jac ? 3 : 6;
  if (o.kind == P_ICHOL)
    per_iter += tri_launches(*o.ic->lplan, &o.ic->lell, false) +
                tri_launches(*o.ic->uplan, &o.ic->uell, false);
  if (o.kind == P_GS_STD)
    per_iter += (tri_launches(*o.Lt, &o.lell, false) + tri_launches(*o.Ut, &o.uell, false) + 4) *
                o.k;
  if (o.kind == P_IT_JACOBI || o.kind == P_GS_IT || o.kind == P_GS_IT2)
    per_iter += 8 * o.m * (o.k + 2);
  int G = std::max(1, std::min(options().graph_iters, 4096 / per_iter));
  G = std::min(G, maxit + 2);

  int status = LAP_OK;
  int *flags = pinned_flags();
  if (!flags) status = LAP_ERR_HIP;
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
  if (status == LAP_OK) {
    hipEvent_t pe[2];
    LAP_HIP(hipEventCreateWithFlags(&pe[0], hipEventDisableTiming));
    LAP_HIP(hipEventCreateWithFlags(&pe[1], hipEventDisableTiming));
    // iter <= maxit + 1, so this many batches always reach done
    const long max_batches = (long)(maxit + 1) / G + 3;
    bool finished = false;
    for (long k = 0; k < max_batches + 1 && !finished; ++k) {
      if (k < max_batches) {
        if (gexec) {
          if (!LAP_HIP(hipGraphLaunch(gexec, stream()))) break;
        } else {
          for (int s = 0; s < G; ++s) iteration();
        }
        LAP_HIP(hipMemcpyAsync(flags + (k & 1), poll, sizeof(int), hipMemcpyDeviceToHost,
                               stream()));
        LAP_HIP(hipEventRecord(pe[k & 1], stream()));
      }
      if (k > 0) {  // inspect the previous batch while this one runs
        LAP_HIP(hipEventSynchronize(pe[(k - 1) & 1]));
        if (flags[(k - 1) & 1]) finished = true;
      }
    }
    if (!finished) {
      set_error(LAP_ERR_NUMERIC, "pcg: convergence flag never set");
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

  PcgCtl h{};
  LAP_HIP(hipMemcpy(&h, ctl, sizeof h, hipMemcpyDeviceToHost));
  out.iterations = h.last_iter;
  out.flag = h.flag;
  out.res0 = h.res0;
  out.res = h.res;
  out.gpu_ms = ms;
  if (h_hist) LAP_HIP(hipMemcpy(h_hist, hist, sizeof(double) * (size_t)(h.last_iter + 1),
                               hipMemcpyDeviceToHost));
  if (status == LAP_OK && lap_last_status() == LAP_ERR_HIP) status = LAP_ERR_HIP;
  if (chunked && tri_error()) {
    set_error(LAP_ERR_NUMERIC, "pcg: a triangular-solve dependency wait ran out");
    tri_clear_error();
    if (status == LAP_OK) status = LAP_ERR_NUMERIC;
  }
  return status;
}

// ||b - A x|| and ||b|| for the final report (cg.c:112-117)
static void true_residual(int n, int nnz, const int *ia, const int *ja, const double *a,
                          const double *x, const double *b, double *tr, double *bn) {
  const SpmvPlan *A = spmv_plan(n, nnz, ia, ja);
  double *r = scratch_vec("pcg_r", (size_t)n);
  spmv_launch(*A, ia, ja, a, x, r, 1.0, 0.0);
  launch_axpy(n, -1.0, b, r);
  ReduceWs &ws = reduce_ws();
  launch_dot(n, r, r, ws.scalars + 1);
  launch_dot(n, b, b, ws.scalars + 2);
  double h[2];
  LAP_HIP(hipMemcpyAsync(h, ws.scalars + 1, 2 * sizeof(double), hipMemcpyDeviceToHost,
                         stream()));
  sync();
  *tr = sqrt(h[0]);
  *bn = sqrt(h[1]);
}

}  // namespace lap

using namespace lap;

extern "C" {

int lap_pcg(int n, int nnz, const int *ia, const int *ja, const double *a, double *x,
            const double *b, double tol, lap_pdata *prec_data, int maxit, double *history,
            lap_solve_info *info) {
  SolveOut out;
  const int st = run_pcg(n, nnz, ia, ja, a, x, b, tol, prec_data, maxit, history, out);
  if (info) {
    info->iterations = out.iterations;
    info->flag = st == LAP_OK ? out.flag : 2;
    info->res0 = out.res0;
    info->res_final = out.res;
    info->gpu_ms = out.gpu_ms;
    info->true_res = -1.0;
    if (st == LAP_OK) {
      double tr, bn;
      true_residual(n, nnz, ia, ja, a, x, b, &tr, &bn);
      info->true_res = tr;
    }
  }
  return st;
}

void cg(int n, double nnz, int *ia, int *ja, double *a, double *x, double *b, double tol,
        lap_pdata *prec_data, int maxit, int *it, int *flag, double *res_norm_history) {
  SolveOut out;
  const int nz = (int)nnz;
  const int st = run_pcg(n, nz, ia, ja, a, x, b, tol, prec_data, maxit, res_norm_history, out);
  *it = out.iterations;
  *flag = st == LAP_OK ? out.flag : 2;
  if (st != LAP_OK) return;
  for (int i = 0; i <= out.iterations; ++i)
    printf("CG: it %d, res norm %5.5e \n", i, res_norm_history[i]);
  if (out.flag == 0) {
    double tr, bn;
    true_residual(n, nz, ia, ja, a, x, b, &tr, &bn);
    printf("TRUE Norm of r %5.5e relative %16.16e\n", tr, tr / bn);
  }
}

double lap_pcg_spmv_timed(int n, int nnz, const int *ia, const int *ja, const double *a,
                          const double *p, double *q, int reps) {
  if (n <= 0 || reps < 1 ||
      !check_device("lap_pcg_spmv_timed", {{ia, "ia"}, {ja, "ja"}, {a, "a"}, {p, "p"}, {q, "q"}}))
    return -1.0;
  const SpmvPlan *A = spmv_plan(n, nnz, ia, ja);
  if (!A) return -1.0;
  ReduceWs &ws = reduce_ws();
  PcgCtl h{};
  h.rho = 1.0;
  PcgCtl *ctl = reinterpret_cast<PcgCtl *>(scratch_vec("timed_ctl", sizeof(PcgCtl) / 8 + 1));
  LAP_HIP(hipMemcpyAsync(ctl, &h, sizeof h, hipMemcpyHostToDevice, stream()));
  const unsigned g = spmv_grid(*A);
  const ViVals vv = vi_prepare(nnz, a, n, ia, ja);  // the kernel the solve would run
  auto launch = [&]() { launch_spmv_pq(g, *A, ia, ja, a, p, q, ctl, ws.partials + kR2, vv); };
  launch();
  hipEvent_t e0, e1;
  LAP_HIP(hipEventCreate(&e0));
  LAP_HIP(hipEventCreate(&e1));
  LAP_HIP(hipEventRecord(e0, stream()));
  for (int i = 0; i < reps; ++i) launch();
  LAP_HIP(hipEventRecord(e1, stream()));
  LAP_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  LAP_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / reps;
}

void prec_function(int *ia, int *ja, double *a, int nnzA, lap_pdata *prec_data, double *x,
                   double *y) {
  PrecOps o;
  if (!prec_data || prec_data->n <= 0) {
    set_error(LAP_ERR_ARG, "prec_function: missing prec_data");
    return;
  }
  if (!check_device("prec_function",
                    {{ia, "ia"}, {ja, "ja"}, {a, "a"}, {x, "x"}, {y, "y"}}))
    return;
  if (!prec_setup(o, prec_data->n, nnzA, ia, ja, a, prec_data)) return;
  prec_launch(o, x, y, nullptr);
  sync();
  if (options().sptrsv_mode != 0 && tri_error()) {
    set_error(LAP_ERR_NUMERIC, "prec_function: a triangular-solve dependency wait ran out");
    tri_clear_error();
  }
}

static void prec_named(const char *name, int *ia, int *ja, double *a, int nnzA,
                       lap_pdata *prec_data, double *in, double *out) {
  lap_pdata tmp = *prec_data;
  tmp.prec_op = const_cast<char *>(name);
  prec_function(ia, ja, a, nnzA, &tmp, in, out);
}

void GS_std(int *ia, int *ja, double *a, int nnzA, lap_pdata *prec_data, double *vec_in,
            double *vec_out) {
  prec_named("GS_std", ia, ja, a, nnzA, prec_data, vec_in, vec_out);
}
void GS_it(int *ia, int *ja, double *a, int nnzA, lap_pdata *prec_data, double *vec_in,
           double *vec_out) {
  prec_named("GS_it", ia, ja, a, nnzA, prec_data, vec_in, vec_out);
}
void GS_it2(int *ia, int *ja, double *a, int nnzA, lap_pdata *prec_data, double *vec_in,
            double *vec_out) {
  prec_named("GS_it2", ia, ja, a, nnzA, prec_data, vec_in, vec_out);
}
void it_jacobi(int *ia, int *ja, double *a, int nnzA, lap_pdata *prec_data, double *vec_in,
               double *vec_out) {
  prec_named("it_jacobi", ia, ja, a, nnzA, prec_data, vec_in, vec_out);
}
void line_jacobi(int *ia, int *ja, double *a, int nnzA, lap_pdata *prec_data, double *vec_in,
                 double *vec_out) {
  prec_named("line_jacobi", ia, ja, a, nnzA, prec_data, vec_in, vec_out);
}

}  // extern "C"
