// batched.hip -- many independent SPD systems solved by Jacobi PCG in one
// persistent launch (lap_pcg_batched).
//
// Layout: the systems are concatenated into one block-diagonal CSR (global
// row and column indices, rows of system s in [row_off[s], row_off[s+1])),
// with concatenated x, b and 1/diag.
//
// Execution: a persistent grid; each workgroup takes the next system from an
// atomic work queue (systems ordered by size, largest first) and runs that
// system's whole PCG loop (cg.c:5-140 with line_jacobi) with workgroup-level
// reductions only, then takes the next.  No workgroup ever waits for another,
// so every wave exits once the queue is drained; uneven sizes are balanced
// by the queue.  Threads per row in the SpMV and the workgroup size are
// chosen per batch by a measured autotuner (lap_pcg_batched_tune).  With one
// thread per row every row is summed in CSR order, as the reference does.
//
// Lockstep mode (block = -1): a batch whose largest system dominates is
// bounded by that system's time on one CU in the persistent kernel.  In
// lockstep every system iterates together over the block-diagonal matrix:
// row blocks never cross a system boundary, each iteration is five launches
// (SpMV with per-block p.q partials; one workgroup computing alpha per
// system; x/r update with per-block partials; one workgroup doing every
// system's convergence test and beta; p update) replayed from a HIP graph,
// and the row blocks of systems that have stopped are skipped.  The
// autotuner times it next to the persistent configurations.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "lap_core.hpp"
#include "spmv_core.hpp"

namespace lap {

namespace {

template <int BS>
__device__ __forceinline__ double block_sum_all(double v, double *red) {
  const double t = block_sum<BS>(v, red);
  if (threadIdx.x == 0) red[0] = t;
  __syncthreads();
  const double r = red[0];
  __syncthreads();
  return r;
}

// Rows [r0, r1) of y = A x, TPR lanes per row (TPR divides 64), handing each
// row's sum to f(row, s) in lane 0 of its group.
template <int BS, int TPR, class F>
__device__ __forceinline__ void rows_spmv(int r0, int r1, const int *__restrict__ ia,
                                          const int *__restrict__ ja,
                                          const double *__restrict__ a,
                                          const double *__restrict__ x, F &&f) {
  const int lane = (int)threadIdx.x % TPR, grp = (int)threadIdx.x / TPR;
  constexpr int NG = BS / TPR;
  for (int row = r0 + grp; row < r1; row += NG) {
    double s = 0.0;
    const int e1 = ia[row + 1];
    for (int e = ia[row] + lane; e < e1; e += TPR) s += a[e] * x[ja[e]];
    if constexpr (TPR > 1) {
#pragma unroll
      for (int o = TPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, TPR);
    }
    if (lane == 0) f(row, s);
  }
}

template <int BS, int TPR>
__global__ __launch_bounds__(BS) void k_batched(
    int nsys, const int *__restrict__ order, const int *__restrict__ row_off,
    const int *__restrict__ ia, const int *__restrict__ ja, const double *__restrict__ a,
    const double *__restrict__ d_r, double *__restrict__ x, const double *__restrict__ b,
    double *__restrict__ r, double *__restrict__ p, double *__restrict__ q, double tol,
    int maxit, int *queue, int *it_out, int *flag_out, double *res0_out, double *res_out) {
  __shared__ double red[BS / kWave];
  __shared__ int slot;
  for (;;) {
    if (threadIdx.x == 0) slot = atomicAdd(queue, 1);
    __syncthreads();
    // slot and the stop flag below are workgroup-uniform; readfirstlane
    // makes every branch on them a scalar branch, so no barrier of this loop
    // can ever be reached by part of a wave
    const int k = __builtin_amdgcn_readfirstlane(slot);
    __syncthreads();
    if (k >= nsys) return;
    const int sys = order[k];
    const int r0 = row_off[sys], r1 = row_off[sys + 1];

    // r = -(A x - b) (cg.c:48-55)
    double acc = 0.0;
    rows_spmv<BS, TPR>(r0, r1, ia, ja, a, x, [&](int row, double s) {
      const double v = -(s - b[row]);
      r[row] = v;
      acc += v * v;
    });
    const double res0 = sqrt(block_sum_all<BS>(acc, red));
    int it = 0, flag = -1;
    double res = res0;
    if (!isfinite(res0)) flag = 2;
    else if (res0 == 0.0) flag = 0;
    flag = __builtin_amdgcn_readfirstlane(flag);
    const double tolrel = tol * res0;
    // w = d_r r, rho = r.w, p = w
    double rho = 0.0;
    if (flag < 0) {
      acc = 0.0;
      for (int i = r0 + (int)threadIdx.x; i < r1; i += BS) {
        const double ri = r[i];
        const double wi = d_r[i] * ri;
        p[i] = wi;
        acc += ri * wi;
      }
      rho = block_sum_all<BS>(acc, red);  // barrier: p visible to the SpMV
    }
    while (flag < 0) {
      // q = A p, p.q
      acc = 0.0;
      rows_spmv<BS, TPR>(r0, r1, ia, ja, a, p, [&](int row, double s) {
        q[row] = s;
        acc += p[row] * s;
      });
      const double pq = block_sum_all<BS>(acc, red);
      const double alpha = rho / pq, nalpha = -1.0 * alpha;
      // x += alpha p; r -= alpha q; r.w and r.r
      double rz = 0.0, rr = 0.0;
      for (int i = r0 + (int)threadIdx.x; i < r1; i += BS) {
        x[i] += alpha * p[i];
        const double ri = r[i] + nalpha * q[i];
        r[i] = ri;
        rz += ri * (d_r[i] * ri);
        rr += ri * ri;
      }
      rz = block_sum_all<BS>(rz, red);
      rr = block_sum_all<BS>(rr, red);
      ++it;
      res = sqrt(rr);
      if (!isfinite(res)) flag = 2;
      else if (res < tolrel) flag = 0;
      else if (it > maxit) flag = 1;
      flag = __builtin_amdgcn_readfirstlane(flag);
      if (flag >= 0) break;
      const double beta = rz / rho;
      rho = rz;
      for (int i = r0 + (int)threadIdx.x; i < r1; i += BS) p[i] = p[i] * beta + d_r[i] * r[i];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      it_out[sys] = it;
      flag_out[sys] = flag;
      res0_out[sys] = res0;
      res_out[sys] = res;
    }
    // Keeps the thread-0 write above and the thread-0 queue pop of the next
    // round in separate barrier intervals: without it the compiler may
    // thread the two thread-0 branches into an inner loop that only lane 0
    // leaves, sending the other lanes back to the slot read (a hang).
    __syncthreads();
  }
}


// ------------------------------------------------------------ lockstep

// Per-system state of the lockstep mode (arrays of nsys).
struct LsSys {
  double *rho, *alpha, *beta, *tolrel, *res, *res0;
  int *it, *flag, *done;
};
struct LsCtl {
  int iter, maxit, all_done, pad;
};

constexpr int kLsRed = 1024;  // one-workgroup per-system reductions

// Sum of part[b*stride] over blocks [b0, b1) in a fixed order, returned to
// every lane of the wave.
__device__ __forceinline__ double ls_sys_sum(const double *part, int stride, int b0, int b1) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  for (int b = b0 + lane; b < b1; b += 64) s += part[(size_t)b * stride];
  return wave_sum(s);
}

struct LsInitEpi {
  double *r;
  const double *rhs;
  double acc;
  __device__ double init(int) const { return 0.0; }
  __device__ void store(int row, double v) {
    const double t = -(v - rhs[row]);  // r = A x; r += -1*b; r *= -1
    r[row] = t;
    acc += t * t;
  }
};

struct LsPqEpi {
  double *q;
  const double *p;
  double acc;
  __device__ double init(int) const { return 0.0; }
  __device__ void store(int row, double v) {
    st_nt(q + row, v);  // q is read once, by the update
    acc += p[row] * v;
  }
};

// r = -(A x - b), per-block partials of r.r
template <int BS, int TILE>
__global__ __launch_bounds__(BS) void k_ls_init(int nblk, const int *__restrict__ rowblk,
                                                const int *__restrict__ ia,
                                                const int *__restrict__ ja,
                                                const double *__restrict__ a,
                                                const double *__restrict__ x,
                                                const double *__restrict__ b,
                                                double *__restrict__ r, double *part) {
  __shared__ double prod[TILE];
  __shared__ double red[BS / kWave];
  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    LsInitEpi epi{r, b, 0.0};
    spmv_block_rows<BS, TILE>(blk, rowblk, ia, ja, a, x, 1.0, prod, red, epi);
    const double t = block_sum<BS>(epi.acc, red);
    if (threadIdx.x == 0) part[blk] = t;
  }
}

// ||r0|| per system, tolrel, stop state; ctl
__global__ __launch_bounds__(kLsRed) void k_ls_start(int nsys, const int *__restrict__ sblk,
                                                     const double *part, double tol, int maxit,
                                                     LsSys S, LsCtl *ctl) {
  __shared__ int live[kLsRed / 64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int cnt = 0;
  for (int s = wave; s < nsys; s += kLsRed / 64) {
    const double r0 = sqrt(ls_sys_sum(part, 1, sblk[s], sblk[s + 1]));
    if (lane == 0) {
      const bool stop = !isfinite(r0) || r0 == 0.0;  // r0 == 0: x0 solves system s
      S.res0[s] = r0;
      S.res[s] = r0;
      S.tolrel[s] = tol * r0;
      S.it[s] = 0;
      S.flag[s] = isfinite(r0) ? 0 : 2;
      S.done[s] = stop ? 1 : 0;
      cnt += stop ? 0 : 1;
    }
  }
  if (lane == 0) live[wave] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kLsRed / 64; ++w) t += live[w];
    ctl->iter = 0;
    ctl->maxit = maxit;
    ctl->all_done = t ? 0 : 1;
  }
}

// Block-uniform "system of this block has stopped" test.
__device__ __forceinline__ bool ls_skip(int blk, const int *__restrict__ bsys,
                                        const int *__restrict__ done) {
  return __builtin_amdgcn_readfirstlane(done[bsys[blk]]) != 0;
}

// w = d_r r, p = w (first iteration), per-block partials of r.w
template <int BS>
__global__ __launch_bounds__(BS) void k_ls_jstart(int nblk, const int *__restrict__ rowblk,
                                                  const int *__restrict__ bsys, const int *done,
                                                  const double *__restrict__ r,
                                                  const double *__restrict__ d_r,
                                                  double *__restrict__ p, double *part) {
  __shared__ double red[BS / kWave];
  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    if (ls_skip(blk, bsys, done)) continue;
    const int r0 = rowblk[blk], r1 = rowblk[blk + 1];
    double acc = 0.0;
    for (int i = r0 + (int)threadIdx.x; i < r1; i += BS) {
      const double ri = r[i];
      const double wi = d_r[i] * ri;  // line_jacobi: w = d_r .* r
      p[i] = wi;
      acc += ri * wi;
    }
    const double t = block_sum<BS>(acc, red);
    if (threadIdx.x == 0) part[blk] = t;
  }
}

// rho per live system
__global__ __launch_bounds__(kLsRed) void k_ls_rho(int nsys, const int *__restrict__ sblk,
                                                   const double *part, LsSys S) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int s = wave; s < nsys; s += kLsRed / 64) {
    if (S.done[s]) continue;
    const double rho = ls_sys_sum(part, 1, sblk[s], sblk[s + 1]);
    if (lane == 0) S.rho[s] = rho;
  }
}

// q = A p for the live systems' blocks, per-block partials of p.q
template <int BS, int TILE>
__global__ __launch_bounds__(BS) void k_ls_spmv(int nblk, const int *__restrict__ rowblk,
                                                const int *__restrict__ bsys, const int *done,
                                                const int *__restrict__ ia,
                                                const int *__restrict__ ja,
                                                const double *__restrict__ a,
                                                const double *__restrict__ p,
                                                double *__restrict__ q, double *part,
                                                const LsCtl *ctl) {
  if (__builtin_amdgcn_readfirstlane(ctl->all_done)) return;
  __shared__ double prod[TILE];
  __shared__ double red[BS / kWave];
  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    if (ls_skip(blk, bsys, done)) continue;
    LsPqEpi epi{q, p, 0.0};
    spmv_block_rows<BS, TILE>(blk, rowblk, ia, ja, a, p, 1.0, prod, red, epi);
    const double t = block_sum<BS>(epi.acc, red);
    if (threadIdx.x == 0) part[blk] = t;
  }
}

// alpha = rho / p.q per live system
__global__ __launch_bounds__(kLsRed) void k_ls_alpha(int nsys, const int *__restrict__ sblk,
                                                     const double *part, LsSys S,
                                                     const LsCtl *ctl) {
  if (ctl->all_done) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int s = wave; s < nsys; s += kLsRed / 64) {
    if (S.done[s]) continue;
    const double pq = ls_sys_sum(part, 1, sblk[s], sblk[s + 1]);
    if (lane == 0) S.alpha[s] = S.rho[s] / pq;
  }
}

// x += alpha p; r -= alpha q; per-block partials of r.w and r.r
template <int BS>
__global__ __launch_bounds__(BS) void k_ls_update(int nblk, const int *__restrict__ rowblk,
                                                  const int *__restrict__ bsys, LsSys S,
                                                  double *__restrict__ x,
                                                  double *__restrict__ r,
                                                  const double *__restrict__ p,
                                                  const double *__restrict__ q,
                                                  const double *__restrict__ d_r, double *part,
                                                  const LsCtl *ctl) {
  if (__builtin_amdgcn_readfirstlane(ctl->all_done)) return;
  __shared__ double red[BS / kWave];
  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    if (ls_skip(blk, bsys, S.done)) continue;
    const int sys = bsys[blk];
    const double alpha = S.alpha[sys], nalpha = -1.0 * alpha;
    const int r0 = rowblk[blk], r1 = rowblk[blk + 1];
    double rz = 0.0, rr = 0.0;
    for (int i = r0 + (int)threadIdx.x; i < r1; i += BS) {  // streamed: non-temporal
      st_nt(x + i, ld_nt(x + i) + alpha * ld_nt(p + i));
      const double ri = ld_nt(r + i) + nalpha * ld_nt(q + i);
      st_nt(r + i, ri);
      rz += ri * (ld_nt(d_r + i) * ri);
      rr += ri * ri;
    }
    const double t_rz = block_sum<BS>(rz, red);
    const double t_rr = block_sum<BS>(rr, red);
    if (threadIdx.x == 0) {
      part[2 * (size_t)blk] = t_rz;
      part[2 * (size_t)blk + 1] = t_rr;
    }
  }
}

// per-system convergence test (cg.c:106-123) and beta; all_done
__global__ __launch_bounds__(kLsRed) void k_ls_conv(int nsys, const int *__restrict__ sblk,
                                                    const double *part, LsSys S, LsCtl *ctl) {
  if (ctl->all_done) return;
  __shared__ int live[kLsRed / 64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int maxit = ctl->maxit;
  int cnt = 0;
  for (int s = wave; s < nsys; s += kLsRed / 64) {
    if (S.done[s]) continue;
    const double rz = ls_sys_sum(part, 2, sblk[s], sblk[s + 1]);
    const double rr = ls_sys_sum(part + 1, 2, sblk[s], sblk[s + 1]);
    if (lane == 0) {
      const int it = S.it[s] + 1;
      const double res = sqrt(rr);
      S.it[s] = it;
      S.res[s] = res;
      int flag = -1;
      if (!isfinite(res))
        flag = 2;
      else if (res < S.tolrel[s])
        flag = 0;
      else if (it > maxit)
        flag = 1;
      if (flag >= 0) {
        S.flag[s] = flag;
        S.done[s] = 1;
      } else {
        S.beta[s] = rz / S.rho[s];
        S.rho[s] = rz;
        ++cnt;
      }
    }
  }
  if (lane == 0) live[wave] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kLsRed / 64; ++w) t += live[w];
    ctl->iter += 1;
    ctl->all_done = t ? 0 : 1;
  }
}

// p = beta p + d_r r for the live systems
template <int BS>
__global__ __launch_bounds__(BS) void k_ls_p(int nblk, const int *__restrict__ rowblk,
                                             const int *__restrict__ bsys, LsSys S,
                                             const double *__restrict__ r,
                                             const double *__restrict__ d_r,
                                             double *__restrict__ p, const LsCtl *ctl) {
  if (__builtin_amdgcn_readfirstlane(ctl->all_done)) return;
  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    if (ls_skip(blk, bsys, S.done)) continue;
    const double beta = S.beta[bsys[blk]];
    const int r0 = rowblk[blk], r1 = rowblk[blk + 1];
    for (int i = r0 + (int)threadIdx.x; i < r1; i += BS)
      st_nt(p + i, ld_nt(p + i) * beta + ld_nt(d_r + i) * ld_nt(r + i));
  }
}

// Row blocks per system: at most kLsRows rows and kSpmvTile nonzeros per
// block (a longer row gets a block of its own), never crossing systems.
constexpr int kLsRows = 256;

struct LsPlan {
  int nblk = 0;
  DevBufPtr rowblk, bsys, sblk, sys;  // sys: LsSys arrays + LsCtl
  LsSys S{};
  LsCtl *ctl = nullptr;
};

struct BatchCfg {
  int bs = 256, tpr = 1;
};

using Launcher = void (*)(unsigned, int, const int *, const int *, const int *, const int *,
                          const double *, const double *, double *, const double *, double *,
                          double *, double *, double, int, int *, int *, int *, double *,
                          double *);

template <int BS, int TPR>
void launch_cfg(unsigned g, int nsys, const int *order, const int *row_off, const int *ia,
                const int *ja, const double *a, const double *d_r, double *x, const double *b,
                double *r, double *p, double *q, double tol, int maxit, int *queue, int *it_out,
                int *flag_out, double *res0_out, double *res_out) {
  k_batched<BS, TPR><<<g, BS, 0, stream()>>>(nsys, order, row_off, ia, ja, a, d_r, x, b, r, p,
                                             q, tol, maxit, queue, it_out, flag_out, res0_out,
                                             res_out);
}

struct CfgEntry {
  int bs, tpr;
  Launcher fn;
};
const CfgEntry kCfgs[] = {
    {256, 1, launch_cfg<256, 1>},   {256, 2, launch_cfg<256, 2>},   {256, 4, launch_cfg<256, 4>},
    {512, 1, launch_cfg<512, 1>},   {512, 2, launch_cfg<512, 2>},   {512, 4, launch_cfg<512, 4>},
    {1024, 1, launch_cfg<1024, 1>}, {1024, 2, launch_cfg<1024, 2>}, {1024, 4, launch_cfg<1024, 4>},
};

const CfgEntry *find_cfg(int bs, int tpr) {
  for (const CfgEntry &c : kCfgs)
    if (c.bs == bs && c.tpr == tpr) return &c;
  return nullptr;
}

// Per-batch state: LPT order, output arrays, scratch vectors.
struct Batch {
  int nsys = 0, n = 0;
  DevBufPtr order, it_out, flag_out, res0_out, res_out, queue;
  int tuned_bs = 0, tuned_tpr = 0;
  std::unique_ptr<LsPlan> ls;
  std::vector<int> off;  // host copy of row_off
};
using BKey = std::tuple<const int *, int, int>;
std::map<BKey, std::unique_ptr<Batch>> g_batches;

Batch *get_batch(int nsys, const int *row_off, int n_total) {
  BKey key{row_off, nsys, n_total};
  auto it = g_batches.find(key);
  if (it != g_batches.end()) return it->second.get();
  std::vector<int> off = dev_download(row_off, (size_t)nsys + 1);
  if (off[0] != 0 || off[nsys] != n_total) {
    set_error(LAP_ERR_ARG, "pcg_batched: row_off must run from 0 to n_total=%d", n_total);
    return nullptr;
  }
  for (int s = 0; s < nsys; ++s)
    if (off[s + 1] < off[s]) {
      set_error(LAP_ERR_ARG, "pcg_batched: row_off decreases at system %d", s);
      return nullptr;
    }
  std::vector<int> ord((size_t)nsys);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int u, int v) {
    return off[u + 1] - off[u] > off[v + 1] - off[v];
  });
  auto B = std::make_unique<Batch>();
  B->nsys = nsys;
  B->n = n_total;
  B->off = off;
  B->order = dev_upload(ord);
  B->it_out = dev_alloc(sizeof(int) * (size_t)nsys);
  B->flag_out = dev_alloc(sizeof(int) * (size_t)nsys);
  B->res0_out = dev_alloc(sizeof(double) * (size_t)nsys);
  B->res_out = dev_alloc(sizeof(double) * (size_t)nsys);
  B->queue = dev_alloc(sizeof(int));
  Batch *raw = B.get();
  g_batches[key] = std::move(B);
  return raw;
}

unsigned batch_grid(int nsys, int bs) {
  const int per_cu = std::max(1, 2048 / bs);  // 32 waves per CU
  return (unsigned)std::max(1, std::min(nsys, per_cu * device_cus()));
}

float run_cfg(const CfgEntry &c, Batch &B, const int *row_off, const int *ia, const int *ja,
              const double *a, const double *d_r, double *x, const double *b, double tol,
              int maxit) {
  double *r = scratch_vec("bt_r", (size_t)B.n), *p = scratch_vec("bt_p", (size_t)B.n),
         *q = scratch_vec("bt_q", (size_t)B.n);
  hipEvent_t e0, e1;
  LAP_HIP(hipEventCreate(&e0));
  LAP_HIP(hipEventCreate(&e1));
  LAP_HIP(hipMemsetAsync(B.queue->p, 0, sizeof(int), stream()));
  LAP_HIP(hipEventRecord(e0, stream()));
  c.fn(batch_grid(B.nsys, c.bs), B.nsys, B.order->as<int>(), row_off, ia, ja, a, d_r, x, b, r, p,
       q, tol, maxit, B.queue->as<int>(), B.it_out->as<int>(), B.flag_out->as<int>(),
       B.res0_out->as<double>(), B.res_out->as<double>());
  LAP_HIP(hipEventRecord(e1, stream()));
  LAP_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  LAP_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}


// Lockstep plan of a batch (host analysis of the row pointers).
LsPlan *ls_plan(Batch &B, const int *ia) {
  if (B.ls) return B.ls.get();
  std::vector<int> h_ia = dev_download(ia, (size_t)B.n + 1);
  std::vector<int> rowblk, bsys, sblk;
  rowblk.push_back(0);
  sblk.push_back(0);
  for (int s = 0; s < B.nsys; ++s) {
    int r = B.off[s];
    const int end = B.off[s + 1];
    while (r < end) {
      int r1 = r + 1;
      if (h_ia[r1] - h_ia[r] <= kSpmvTile)
        while (r1 < end && r1 - r < kLsRows && h_ia[r1 + 1] - h_ia[r] <= kSpmvTile) ++r1;
      rowblk.push_back(r1);
      bsys.push_back(s);
      r = r1;
    }
    sblk.push_back((int)bsys.size());
  }
  auto P = std::make_unique<LsPlan>();
  P->nblk = (int)bsys.size();
  P->rowblk = dev_upload(rowblk);
  P->bsys = dev_upload(bsys.empty() ? std::vector<int>{0} : bsys);
  P->sblk = dev_upload(sblk);
  const size_t ns = (size_t)B.nsys;
  const size_t dbytes = sizeof(double) * ns * 6, ibytes = sizeof(int) * ns * 3;
  P->sys = dev_alloc(dbytes + ibytes + sizeof(LsCtl) + 64);
  if (!P->rowblk || !P->bsys || !P->sblk || !P->sys) return nullptr;
  char *base = P->sys->as<char>();
  double *d = reinterpret_cast<double *>(base);
  P->S = LsSys{d, d + ns, d + 2 * ns, d + 3 * ns, d + 4 * ns, d + 5 * ns, nullptr, nullptr,
               nullptr};
  int *iv = reinterpret_cast<int *>(base + dbytes);
  P->S.it = iv;
  P->S.flag = iv + ns;
  P->S.done = iv + 2 * ns;
  P->ctl = reinterpret_cast<LsCtl *>(base + ((dbytes + ibytes + 15) & ~(size_t)15));
  B.ls = std::move(P);
  return B.ls.get();
}

// One lockstep solve; outputs land in B's output arrays.  Returns the
// device milliseconds (HIP events), or -1 on error.
float run_lockstep(Batch &B, const int *ia, const int *ja, const double *a, const double *d_r,
                   double *x, const double *b, double tol, int maxit) {
  LsPlan *P = ls_plan(B, ia);
  if (!P) return -1.f;
  const int nblk = P->nblk;
  double *r = scratch_vec("bt_r", (size_t)B.n), *p = scratch_vec("bt_p", (size_t)B.n),
         *q = scratch_vec("bt_q", (size_t)B.n);
  double *part = scratch_vec("ls_part", (size_t)5 * std::max(nblk, 1));
  if (!r || !p || !q || !part) return -1.f;
  double *part0 = part, *part1 = part + nblk, *part2 = part + 2 * (size_t)nblk,
         *part3 = part + 3 * (size_t)nblk;
  const int *rowblk = P->rowblk->as<int>(), *bsys = P->bsys->as<int>(),
            *sblk = P->sblk->as<int>();
  const unsigned g = (unsigned)std::max(1, std::min(nblk, 8 * device_cus()));
  hipEvent_t e0, e1;
  if (!LAP_HIP(hipEventCreate(&e0))) return -1.f;
  if (!LAP_HIP(hipEventCreate(&e1))) {
    (void)hipEventDestroy(e0);
    return -1.f;
  }
  LAP_HIP(hipEventRecord(e0, stream()));
  k_ls_init<kBlock, kSpmvTile><<<g, kBlock, 0, stream()>>>(nblk, rowblk, ia, ja, a, x, b, r,
                                                          part0);
  k_ls_start<<<1, kLsRed, 0, stream()>>>(B.nsys, sblk, part0, tol, maxit, P->S, P->ctl);
  k_ls_jstart<kBlock><<<g, kBlock, 0, stream()>>>(nblk, rowblk, bsys, P->S.done, r, d_r, p,
                                                  part1);
  k_ls_rho<<<1, kLsRed, 0, stream()>>>(B.nsys, sblk, part1, P->S);
  auto iteration = [&]() {
    k_ls_spmv<kBlock, kSpmvTile><<<g, kBlock, 0, stream()>>>(nblk, rowblk, bsys, P->S.done, ia,
                                                            ja, a, p, q, part2, P->ctl);
    k_ls_alpha<<<1, kLsRed, 0, stream()>>>(B.nsys, sblk, part2, P->S, P->ctl);
    k_ls_update<kBlock><<<g, kBlock, 0, stream()>>>(nblk, rowblk, bsys, P->S, x, r, p, q, d_r,
                                                    part3, P->ctl);
    k_ls_conv<<<1, kLsRed, 0, stream()>>>(B.nsys, sblk, part3, P->S, P->ctl);
    k_ls_p<kBlock><<<g, kBlock, 0, stream()>>>(nblk, rowblk, bsys, P->S, r, d_r, p, P->ctl);
  };
  // every system stops by iteration maxit + 1, so G * max_batches
  // iterations always reach all_done
  const int G = std::max(1, std::min(options().graph_iters, maxit + 2));
  const long max_batches = (long)(maxit + 1) / G + 3;
  bool ok = true;
  hipGraphExec_t gexec = nullptr;
  if (options().use_graph) {
    hipGraph_t graph = nullptr;
    ok = LAP_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
    if (ok) {
      for (int i = 0; i < G; ++i) iteration();
      ok = LAP_HIP(hipStreamEndCapture(stream(), &graph));
    }
    if (ok) ok = LAP_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    if (graph) (void)hipGraphDestroy(graph);
  }
  int *flag_h = pinned_flags();
  bool finished = false;
  if (ok && flag_h) {
    for (long it = 0; it < max_batches && !finished; ++it) {
      if (gexec) {
        if (!LAP_HIP(hipGraphLaunch(gexec, stream()))) break;
      } else {
        for (int i = 0; i < G; ++i) iteration();
      }
      LAP_HIP(hipMemcpyAsync(flag_h, &P->ctl->all_done, sizeof(int), hipMemcpyDeviceToHost,
                             stream()));
      sync();
      finished = flag_h[0] != 0;
    }
  }
  LAP_HIP(hipEventRecord(e1, stream()));
  LAP_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  LAP_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (gexec) (void)hipGraphExecDestroy(gexec);
  const size_t ns = (size_t)B.nsys;
  LAP_HIP(hipMemcpyAsync(B.it_out->p, P->S.it, sizeof(int) * ns, hipMemcpyDeviceToDevice,
                         stream()));
  LAP_HIP(hipMemcpyAsync(B.flag_out->p, P->S.flag, sizeof(int) * ns, hipMemcpyDeviceToDevice,
                         stream()));
  LAP_HIP(hipMemcpyAsync(B.res0_out->p, P->S.res0, sizeof(double) * ns,
                         hipMemcpyDeviceToDevice, stream()));
  LAP_HIP(hipMemcpyAsync(B.res_out->p, P->S.res, sizeof(double) * ns, hipMemcpyDeviceToDevice,
                         stream()));
  sync();
  if (!ok || !finished) {
    if (ok) set_error(LAP_ERR_NUMERIC, "pcg_batched: lockstep stop flags never set");
    return -1.f;
  }
  return ms;
}

bool check_batch_args(int nsys, const int *row_off, int n_total, int nnz_total, const int *ia,
                      const int *ja, const double *a, const double *d_r, const double *x,
                      const double *b) {
  if (nsys <= 0 || n_total <= 0 || nnz_total < 0) {
    set_error(LAP_ERR_ARG, "pcg_batched: bad sizes nsys=%d n=%d nnz=%d", nsys, n_total, nnz_total);
    return false;
  }
  const size_t nd = sizeof(double) * (size_t)n_total;
  if (!check_device("pcg_batched",
                    {{row_off, "row_off", sizeof(int) * ((size_t)nsys + 1)},
                     {ia, "ia", sizeof(int) * ((size_t)n_total + 1)},
                     {ja, "ja", sizeof(int) * (size_t)nnz_total},
                     {a, "a", sizeof(double) * (size_t)nnz_total},
                     {d_r, "d_r", nd}, {x, "x", nd}, {b, "b", nd}}))
    return false;
  // column indices are range-checked by the SpMV analysis of the big CSR
  return spmv_plan(n_total, nnz_total, ia, ja) != nullptr;
}

}  // namespace

void release_batch_plans() { g_batches.clear(); }
void forget_batch_plans(const void *p) {
  for (auto it = g_batches.begin(); it != g_batches.end();) {
    if (std::get<0>(it->first) == p)
      it = g_batches.erase(it);
    else
      ++it;
  }
}

}  // namespace lap

using namespace lap;

extern "C" {

int lap_pcg_batched_tune(int nsys, const int *row_off, int n_total, int nnz_total,
                         const int *ia, const int *ja, const double *a, const double *d_r,
                         const double *b, int trial_iters, int *best_block, int *best_tpr) {
  if (!check_batch_args(nsys, row_off, n_total, nnz_total, ia, ja, a, d_r, b, b))
    return lap_last_status();
  Batch *B = get_batch(nsys, row_off, n_total);
  if (!B) return lap_last_status();
  double *xs = scratch_vec("bt_tune_x", (size_t)n_total);
  float best = 1e30f;
  const CfgEntry *win = &kCfgs[0];
  for (const CfgEntry &c : kCfgs) {
    launch_fill(n_total, 0.0, xs);
    run_cfg(c, *B, row_off, ia, ja, a, d_r, xs, b, 0.0, trial_iters);  // warm
    launch_fill(n_total, 0.0, xs);
    const float ms = run_cfg(c, *B, row_off, ia, ja, a, d_r, xs, b, 0.0, trial_iters);
    if (ms < best) {
      best = ms;
      win = &c;
    }
  }
  int bs_win = win->bs, tpr_win = win->tpr;
  {  // the lockstep mode on the same trial
    launch_fill(n_total, 0.0, xs);
    const float w = run_lockstep(*B, ia, ja, a, d_r, xs, b, 0.0, trial_iters);  // warm
    launch_fill(n_total, 0.0, xs);
    const float ms = w < 0.f ? -1.f : run_lockstep(*B, ia, ja, a, d_r, xs, b, 0.0, trial_iters);
    if (ms < 0.f) return lap_last_status() != LAP_OK ? lap_last_status() : LAP_ERR_HIP;
    if (ms < best) {
      bs_win = -1;
      tpr_win = 0;
    }
  }
  B->tuned_bs = bs_win;
  B->tuned_tpr = tpr_win;
  if (best_block) *best_block = bs_win;
  if (best_tpr) *best_tpr = tpr_win;
  return lap_last_status();
}

int lap_pcg_batched(int nsys, const int *row_off, int n_total, int nnz_total, const int *ia,
                    const int *ja, const double *a, const double *d_r, double *x,
                    const double *b, double tol, int maxit, int block, int tpr, int *iters,
                    int *flags, double *res0, double *res, double *gpu_ms) {
  if (maxit < 0) {
    set_error(LAP_ERR_ARG, "pcg_batched: maxit < 0");
    return LAP_ERR_ARG;
  }
  if (!check_batch_args(nsys, row_off, n_total, nnz_total, ia, ja, a, d_r, x, b))
    return lap_last_status();
  Batch *B = get_batch(nsys, row_off, n_total);
  if (!B) return lap_last_status();
  if (block == 0 && B->tuned_bs) {
    block = B->tuned_bs;
    tpr = B->tuned_tpr;
  }
  if (block == 0) block = 256;
  if (tpr == 0) tpr = 1;
  float ms = 0.f;
  if (block == -1) {
    ms = run_lockstep(*B, ia, ja, a, d_r, x, b, tol, maxit);
    if (ms < 0.f) return lap_last_status() != LAP_OK ? lap_last_status() : LAP_ERR_HIP;
  } else {
    const CfgEntry *c = find_cfg(block, tpr);
    if (!c) {
      set_error(LAP_ERR_ARG, "pcg_batched: no kernel for block=%d tpr=%d", block, tpr);
      return LAP_ERR_ARG;
    }
    ms = run_cfg(*c, *B, row_off, ia, ja, a, d_r, x, b, tol, maxit);
  }
  if (gpu_ms) *gpu_ms = ms;
  sync();
  if (iters) LAP_HIP(hipMemcpy(iters, B->it_out->p, sizeof(int) * (size_t)nsys,
                               hipMemcpyDeviceToHost));
  if (flags) LAP_HIP(hipMemcpy(flags, B->flag_out->p, sizeof(int) * (size_t)nsys,
                               hipMemcpyDeviceToHost));
  if (res0) LAP_HIP(hipMemcpy(res0, B->res0_out->p, sizeof(double) * (size_t)nsys,
                              hipMemcpyDeviceToHost));
  if (res) LAP_HIP(hipMemcpy(res, B->res_out->p, sizeof(double) * (size_t)nsys,
                             hipMemcpyDeviceToHost));
  return lap_last_status();
}

}  // extern "C"
