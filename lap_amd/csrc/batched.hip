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
#include <algorithm>
#include <cmath>
#include <numeric>

#include "lap_core.hpp"

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
  B->tuned_bs = win->bs;
  B->tuned_tpr = win->tpr;
  if (best_block) *best_block = win->bs;
  if (best_tpr) *best_tpr = win->tpr;
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
  const CfgEntry *c = find_cfg(block, tpr);
  if (!c) {
    set_error(LAP_ERR_ARG, "pcg_batched: no kernel for block=%d tpr=%d", block, tpr);
    return LAP_ERR_ARG;
  }
  const float ms = run_cfg(*c, *B, row_off, ia, ja, a, d_r, x, b, tol, maxit);
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
