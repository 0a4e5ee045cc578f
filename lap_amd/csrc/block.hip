// block.hip -- Jacobi PCG for k right-hand sides sharing every SpMM
// (lap_pcg_block), and the SpMM itself (lap_csr_matmat).
//
// The k columns run k independent CG recurrences (cg.c:5-140 each, with
// line_jacobi) whose SpMVs are fused into one SpMM Q = A P: the CSR matrix
// is read once per iteration for all k columns (12 B per nonzero amortised
// over k).  Block vectors are row-interleaved (n x k, element (i, j) at
// i*k + j), so a row of P is one contiguous 8k-byte run and a CSR column
// index addresses a whole run.
//
// SpMM: one column per lane; a wave holds G = 64/k row groups of k lanes.
// Each group takes RU rows per step and, for each row, issues the loads of
// EU nonzeros (column index, value, the P entries) before summing them, so a
// lane keeps RU*EU gathers in flight; the sums themselves run in CSR order,
// one per column, which is the reference's simple_csr_matvec sum
// (simple_blas.c:24-44) bit for bit whatever RU is.  RU (1, 2, 4) is picked
// per (matrix, k) by timing the candidates once.
//
// Per-column scalars (rho, alpha, beta, tolrel, flags) live in a device
// control block.  One iteration is five launches, graph-replayed:
//   spmm    Q = A P, per-workgroup per-column partials of P.Q
//   alpha   one workgroup: alpha_j = rho_j / (p.q)_j
//   update  R -= alpha Q; partials of R.W (W = d_r R) and R.R
//   conv    one workgroup: per-column convergence test (cg.c:106-123), beta
//   dir     X += alpha P for the columns live in this iteration, then
//           P = beta P + d_r R for the columns still live
// Moving x += alpha p from `update` into `dir` reads P once per iteration
// instead of twice; the x of a column that stops is still updated in the
// iteration it stops, as in cg.c (a final launch covers the last iteration,
// whose `dir` is skipped once every column has stopped).  A stopped column
// keeps its x, iteration count and residual, as a single-RHS solve does.
#include <algorithm>
#include <climits>
#include <cmath>

#include "lap_core.hpp"

namespace lap {

namespace {

constexpr int kMaxK = 64;         // columns per block solve
constexpr int kRedThreads = 1024; // one-workgroup column reductions

struct BlkCtl {
  double rho[kMaxK], alpha[kMaxK], beta[kMaxK], tolrel[kMaxK], res0[kMaxK], res[kMaxK];
  int done[kMaxK];  // column stopped (converged / maxit / non-finite)
  int xup[kMaxK];   // column was live in the last finished iteration
  int flag[kMaxK], iters[kMaxK];
  int iter, maxit, all_done, pad;
};

// ---------------------------------------------------------------- SpMM

// Q = A P for K columns.  MODE 0: plain (lap_csr_matmat); 1: partials of
// P.Q per column; 2: Q = -(A P - B) (initial residual) with partials of
// Q.Q.  Partials: part[blockIdx.x * K + col].
template <int BS, int K, int RU, int EU, int MODE>
__global__ __launch_bounds__(BS) void k_spmm(int n, const int *__restrict__ ia,
                                             const int *__restrict__ ja,
                                             const double *__restrict__ a,
                                             const double *__restrict__ P, double *__restrict__ Q,
                                             const double *__restrict__ Bm, double *part,
                                             const BlkCtl *ctl) {
  if (MODE == 1 && ctl->all_done) return;
  constexpr int G = 64 / K, NW = BS / 64;
  __shared__ double lds[K * NW];
  const int lane = threadIdx.x & 63;
  // wave-uniform: with K = 64 every row index below is uniform, so the row
  // pointers, column indices and values are scalar loads
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int g = lane / K, c = lane % K;
  const int wave_id = (int)blockIdx.x * NW + wave, nwaves = (int)gridDim.x * NW;
  double acc = 0.0;
  for (long base = (long)wave_id * G * RU; base < n; base += (long)nwaves * G * RU) {
    int e0[RU], e1[RU];
#pragma unroll
    for (int j = 0; j < RU; ++j) {
      const long row = base + (long)j * G + g;
      e0[j] = row < n ? ia[row] : 0;
      e1[j] = row < n ? ia[row + 1] : 0;
    }
    double s[RU];
#pragma unroll
    for (int j = 0; j < RU; ++j) s[j] = 0.0;
    for (int off = 0;; off += EU) {
      bool any = false;
#pragma unroll
      for (int j = 0; j < RU; ++j) any |= e0[j] + off < e1[j];
      if (!any) break;
      bool ok[RU][EU];
      int col[RU][EU];
      double av[RU][EU], v[RU][EU];
#pragma unroll
      for (int j = 0; j < RU; ++j)
#pragma unroll
        for (int t = 0; t < EU; ++t) {
          const int e = e0[j] + off + t;
          ok[j][t] = e < e1[j];
          col[j][t] = ok[j][t] ? ja[e] : 0;
          av[j][t] = ok[j][t] ? a[e] : 0.0;
        }
#pragma unroll
      for (int j = 0; j < RU; ++j)
#pragma unroll
        for (int t = 0; t < EU; ++t)
          v[j][t] = ok[j][t] ? P[(size_t)col[j][t] * K + c] : 0.0;
#pragma unroll
      for (int j = 0; j < RU; ++j)
#pragma unroll
        for (int t = 0; t < EU; ++t)
          if (ok[j][t]) s[j] += av[j][t] * v[j][t];
    }
#pragma unroll
    for (int j = 0; j < RU; ++j) {
      const long row = base + (long)j * G + g;
      if (row < n) {
        const size_t o = (size_t)row * K + c;
        if (MODE == 2) {
          const double r = -(s[j] - Bm[o]);  // r = A x; r += -1*b; r *= -1
          Q[o] = r;
          acc += r * r;
        } else {
          Q[o] = s[j];
          if (MODE == 1) acc += P[o] * s[j];
        }
      }
    }
  }
  if (MODE == 0) return;
  // fixed-order reduction: butterfly over the row groups of a wave, then
  // over the waves of the workgroup
  for (int off = K; off < 64; off <<= 1) acc += __shfl_xor(acc, off);
  if (g == 0) lds[c * NW + wave] = acc;
  __syncthreads();
  for (int cc = threadIdx.x; cc < K; cc += BS) {
    double t = 0.0;
    for (int w = 0; w < NW; ++w) t += lds[cc * NW + w];
    part[(size_t)blockIdx.x * K + cc] = t;
  }
}

// K = 64, wave-cooperative: one wave = the 64 columns of a row; a step takes
// RU consecutive rows, whose nonzeros are contiguous in CSR, so their
// column indices and values arrive in one coalesced load per 64 entries and
// v_readlane hands each entry to the whole wave.  The gathers of up to 16
// entries are issued before they are summed; rows are summed in CSR order
// and written as soon as they end.  Control flow is wave-uniform.
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int BS, int RU, int MODE>
__global__ __launch_bounds__(BS) void k_spmm64(int n, const int *__restrict__ ia,
                                               const int *__restrict__ ja,
                                               const double *__restrict__ a,
                                               const double *__restrict__ P,
                                               double *__restrict__ Q,
                                               const double *__restrict__ Bm, double *part,
                                               const BlkCtl *ctl) {
  static_assert(RU < 64, "row pointers of a step fit one wave");
  if (MODE == 1 && ctl->all_done) return;
  constexpr int K = 64, NW = BS / 64, EB = 16;
  __shared__ double lds[K * NW];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int wave_id = (int)blockIdx.x * NW + wave, nwaves = (int)gridDim.x * NW;
  double acc = 0.0;
  // P's own row for the p.q epilogue: the diagonal entry's gather already
  // holds it (loaded only for a row without a stored diagonal)
  double pown = 0.0;
  bool have_own = false;
  auto finish = [&](int row, double sum) {
    const size_t o = (size_t)row * K + lane;
    if (MODE == 2) {
      const double r = -(sum - Bm[o]);  // r = A x; r += -1*b; r *= -1
      Q[o] = r;
      acc += r * r;
    } else {
      Q[o] = sum;
      if (MODE == 1) acc += (have_own ? pown : P[o]) * sum;
    }
    have_own = false;
  };
  for (int base = wave_id * RU; base < n; base += nwaves * RU) {
    const int nrows = min(RU, n - base);
    const int ptr_l = lane <= nrows ? ia[base + lane] : 0;
    const int e_begin = __builtin_amdgcn_readlane(ptr_l, 0);
    const int e_end = __builtin_amdgcn_readlane(ptr_l, nrows);
    int j = 0;                                             // row within the step
    int next = __builtin_amdgcn_readlane(ptr_l, 1);        // end of row j
    double cur = 0.0;
    for (int c0 = e_begin; c0 < e_end; c0 += 64) {
      const int cnt = min(64, e_end - c0);
      const int jl = lane < cnt ? ja[c0 + lane] : 0;
      const double al = lane < cnt ? a[c0 + lane] : 0.0;
      for (int t0 = 0; t0 < cnt; t0 += EB) {
        double v[EB];
#pragma unroll
        for (int u = 0; u < EB; ++u)
          if (t0 + u < cnt)
            v[u] = P[(size_t)__builtin_amdgcn_readlane(jl, t0 + u) * K + lane];
#pragma unroll
        for (int u = 0; u < EB; ++u) {
          if (t0 + u < cnt) {
            const int e = c0 + t0 + u;
            while (e >= next) {  // row j ends before entry e
              finish(base + j, cur);
              cur = 0.0;
              ++j;
              next = __builtin_amdgcn_readlane(ptr_l, j + 1);
            }
            if (MODE == 1 && __builtin_amdgcn_readlane(jl, t0 + u) == base + j) {
              pown = v[u];
              have_own = true;
            }
            cur += readlane_f64(al, t0 + u) * v[u];
          }
        }
      }
    }
    while (j < nrows) {  // the last row with entries, then empty rows
      finish(base + j, cur);
      cur = 0.0;
      ++j;
    }
  }
  if (MODE == 0) return;
  lds[lane * NW + wave] = acc;
  __syncthreads();
  for (int cc = threadIdx.x; cc < K; cc += BS) {
    double t = 0.0;
    for (int w = 0; w < NW; ++w) t += lds[cc * NW + w];
    part[(size_t)blockIdx.x * K + cc] = t;
  }
}

// --------------------------------------------------- column reductions

// One workgroup of kRedThreads: out[j] = sum over b < nparts of
// part[b * stride + j], j < k, in a fixed order.
__device__ __forceinline__ void column_sums(const double *part, int nparts, int stride, int k,
                                            double *lds, double *out) {
  const int nchunk = kRedThreads / k;  // k | 1024
  const int col = threadIdx.x % k, chunk = threadIdx.x / k;
  double s = 0.0;
  for (int b = chunk; b < nparts; b += nchunk) s += part[(size_t)b * stride + col];
  lds[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < k) {
    double t = 0.0;
    for (int ch = 0; ch < nchunk; ++ch) t += lds[ch * k + threadIdx.x];
    out[threadIdx.x] = t;
  }
  __syncthreads();
}

// start: ||r0||, tolrel and the stop state of every column
__global__ __launch_bounds__(kRedThreads) void k_blk_res0(const double *part, int nparts, int k,
                                                          double tol, int maxit, BlkCtl *ctl,
                                                          double *hist) {
  __shared__ double lds[kRedThreads];
  __shared__ double sums[kMaxK];
  __shared__ int any_live;
  if (threadIdx.x == 0) any_live = 0;
  column_sums(part, nparts, k, k, lds, sums);
  if (threadIdx.x < k) {
    const int j = threadIdx.x;
    const double r0 = sqrt(sums[j]);
    const bool stop = !isfinite(r0) || r0 == 0.0;  // r0 == 0: x0 solves column j
    ctl->res0[j] = r0;
    ctl->res[j] = r0;
    ctl->tolrel[j] = tol * r0;
    ctl->iters[j] = 0;
    ctl->flag[j] = isfinite(r0) ? 0 : 2;
    ctl->done[j] = stop ? 1 : 0;
    ctl->xup[j] = 0;
    if (hist) hist[j] = r0;
    if (!stop) atomicOr(&any_live, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ctl->iter = 0;
    ctl->maxit = maxit;
    ctl->all_done = any_live ? 0 : 1;
  }
}

// W = d_r R; P = W; per-workgroup partials of R.W
template <int BS>
__global__ __launch_bounds__(BS) void k_blk_start(int n, int k, int lk,
                                                  const double *__restrict__ R,
                                                  const double *__restrict__ d_r,
                                                  double *__restrict__ P, double *part) {
  __shared__ double lds[BS];
  const size_t total = (size_t)n * k, stride = (size_t)gridDim.x * BS;
  double acc = 0.0;  // this thread's column is threadIdx.x % k (k | BS)
  for (size_t o = (size_t)blockIdx.x * BS + threadIdx.x; o < total; o += stride) {
    const double ri = R[o];
    const double wi = d_r[o >> lk] * ri;  // line_jacobi: w = d_r .* r
    P[o] = wi;                            // first iteration: p = w
    acc += ri * wi;
  }
  lds[threadIdx.x] = acc;
  __syncthreads();
  for (int col = threadIdx.x; col < k; col += BS) {
    double s = 0.0;
    for (int t = col; t < BS; t += k) s += lds[t];
    part[(size_t)blockIdx.x * k + col] = s;
  }
}

// rho_j from the R.W partials
__global__ __launch_bounds__(kRedThreads) void k_blk_rho(const double *part, int nparts, int k,
                                                         BlkCtl *ctl) {
  __shared__ double lds[kRedThreads];
  __shared__ double sums[kMaxK];
  column_sums(part, nparts, k, k, lds, sums);
  if (threadIdx.x < k) ctl->rho[threadIdx.x] = sums[threadIdx.x];
}

// alpha_j = rho_j / (p.q)_j
__global__ __launch_bounds__(kRedThreads) void k_blk_alpha(const double *part, int nparts, int k,
                                                           BlkCtl *ctl) {
  if (ctl->all_done) return;
  __shared__ double lds[kRedThreads];
  __shared__ double sums[kMaxK];
  column_sums(part, nparts, k, k, lds, sums);
  if (threadIdx.x < k) ctl->alpha[threadIdx.x] = ctl->rho[threadIdx.x] / sums[threadIdx.x];
}

// R -= alpha Q for the live columns; per-workgroup partials of R.W and R.R
// (part[b*2k + j] and part[b*2k + k + j])
template <int BS>
__global__ __launch_bounds__(BS) void k_blk_update(int n, int k, int lk, double *__restrict__ R,
                                                   const double *__restrict__ Q,
                                                   const double *__restrict__ d_r,
                                                   const BlkCtl *ctl, double *part) {
  if (ctl->all_done) return;
  __shared__ double lds[2 * BS];
  const int col = (int)threadIdx.x & (k - 1);
  const double nalpha = -1.0 * ctl->alpha[col];
  const size_t total = (size_t)n * k, stride = (size_t)gridDim.x * BS;
  double rz = 0.0, rr = 0.0;
  if (!ctl->done[col]) {
    for (size_t o = (size_t)blockIdx.x * BS + threadIdx.x; o < total; o += stride) {
      // R and Q streamed once: non-temporal (d_r, shared by a row's k
      // columns, stays cached)
      const double ri = ld_nt(R + o) + nalpha * ld_nt(Q + o);  // r += -alpha*q
      st_nt(R + o, ri);
      rz += ri * (d_r[o >> lk] * ri);
      rr += ri * ri;
    }
  }
  lds[threadIdx.x] = rz;
  lds[BS + threadIdx.x] = rr;
  __syncthreads();
  for (int cc = threadIdx.x; cc < k; cc += BS) {
    double s1 = 0.0, s2 = 0.0;
    for (int t = cc; t < BS; t += k) {
      s1 += lds[t];
      s2 += lds[BS + t];
    }
    part[(size_t)blockIdx.x * 2 * k + cc] = s1;
    part[(size_t)blockIdx.x * 2 * k + k + cc] = s2;
  }
}

// per-column convergence test (cg.c:106-123) and beta; hist[it*k + j]
__global__ __launch_bounds__(kRedThreads) void k_blk_conv(const double *part, int nparts, int k,
                                                          double *hist, BlkCtl *ctl) {
  if (ctl->all_done) return;
  __shared__ double lds[kRedThreads];
  __shared__ double rz[kMaxK], rr[kMaxK];
  __shared__ int any_live;
  if (threadIdx.x == 0) any_live = 0;
  column_sums(part, nparts, 2 * k, k, lds, rz);
  column_sums(part + k, nparts, 2 * k, k, lds, rr);
  const int it = ctl->iter + 1;
  if (threadIdx.x < k) {
    const int j = threadIdx.x;
    const bool live = !ctl->done[j];
    ctl->xup[j] = live ? 1 : 0;
    if (live) {
      const double res = sqrt(rr[j]);
      ctl->res[j] = res;
      ctl->iters[j] = it;
      if (hist) hist[(size_t)it * k + j] = res;
      int flag = -1;
      if (!isfinite(res))
        flag = 2;
      else if (res < ctl->tolrel[j])
        flag = 0;
      else if (it > ctl->maxit)
        flag = 1;
      if (flag >= 0) {
        ctl->flag[j] = flag;
        ctl->done[j] = 1;
      } else {
        ctl->beta[j] = rz[j] / ctl->rho[j];
        ctl->rho[j] = rz[j];
        atomicOr(&any_live, 1);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ctl->iter = it;
    ctl->all_done = any_live ? 0 : 1;
  }
}

// X += alpha P (columns live in this iteration); P = beta P + d_r R
// (columns still live).  FINAL: the X update of the last iteration only.
template <int BS, bool FINAL>
__global__ __launch_bounds__(BS) void k_blk_dir(int n, int k, int lk, double *__restrict__ X,
                                                const double *__restrict__ R,
                                                const double *__restrict__ d_r,
                                                double *__restrict__ P, const BlkCtl *ctl) {
  if (!FINAL && ctl->all_done) return;
  const int col = (int)threadIdx.x & (k - 1);
  const bool xup = ctl->xup[col], live = !FINAL && !ctl->done[col];
  if (!xup && !live) return;
  const double alpha = ctl->alpha[col], beta = ctl->beta[col];
  const size_t total = (size_t)n * k, stride = (size_t)gridDim.x * BS;
  for (size_t o = (size_t)blockIdx.x * BS + threadIdx.x; o < total; o += stride) {
    const double pv = ld_nt(P + o);
    if (xup) st_nt(X + o, ld_nt(X + o) + alpha * pv);                    // x += alpha*p
    if (live) st_nt(P + o, pv * beta + d_r[o >> lk] * ld_nt(R + o));  // p = beta p + w
  }
}

using SpmmFn = void (*)(unsigned, int, const int *, const int *, const double *, const double *,
                        double *, const double *, double *, const BlkCtl *, int);

template <int K, int RU>
void spmm_launch(unsigned g, int n, const int *ia, const int *ja, const double *a,
                 const double *P, double *Q, const double *Bm, double *part, const BlkCtl *ctl,
                 int mode) {
  constexpr int EU = RU == 1 ? 8 : 4;
  if (mode == 2)
    k_spmm<kBlock, K, RU, EU, 2><<<g, kBlock, 0, stream()>>>(n, ia, ja, a, P, Q, Bm, part, ctl);
  else if (mode == 1)
    k_spmm<kBlock, K, RU, EU, 1><<<g, kBlock, 0, stream()>>>(n, ia, ja, a, P, Q, Bm, part, ctl);
  else
    k_spmm<kBlock, K, RU, EU, 0><<<g, kBlock, 0, stream()>>>(n, ia, ja, a, P, Q, Bm, part, ctl);
}

template <int RU>
void spmm64_launch(unsigned g, int n, const int *ia, const int *ja, const double *a,
                   const double *P, double *Q, const double *Bm, double *part, const BlkCtl *ctl,
                   int mode) {
  if (mode == 2)
    k_spmm64<kBlock, RU, 2><<<g, kBlock, 0, stream()>>>(n, ia, ja, a, P, Q, Bm, part, ctl);
  else if (mode == 1)
    k_spmm64<kBlock, RU, 1><<<g, kBlock, 0, stream()>>>(n, ia, ja, a, P, Q, Bm, part, ctl);
  else
    k_spmm64<kBlock, RU, 0><<<g, kBlock, 0, stream()>>>(n, ia, ja, a, P, Q, Bm, part, ctl);
}

template <int K>
SpmmFn spmm_for_k(int ru) {
  switch (ru) {
    case 1: return spmm_launch<K, 1>;
    case 2: return spmm_launch<K, 2>;
    case 4: return spmm_launch<K, 4>;
    default: return nullptr;
  }
}

SpmmFn spmm_for(int k, int ru) {
  switch (k) {
    case 1: return spmm_for_k<1>(ru);
    case 2: return spmm_for_k<2>(ru);
    case 4: return spmm_for_k<4>(ru);
    case 8: return spmm_for_k<8>(ru);
    case 16: return spmm_for_k<16>(ru);
    case 32: return spmm_for_k<32>(ru);
    case 64:
      if (ru == 8) return spmm64_launch<8>;
      if (ru == 16) return spmm64_launch<16>;
      return spmm_for_k<64>(ru);
    default: return nullptr;
  }
}

// 1, 2, 4: rows per lane group; 8, 16 (k = 64): rows per wave step with
// wave-cooperative index loads
bool tile_ok(int k, int ru) {
  return ru == 1 || ru == 2 || ru == 4 || (k == 64 && (ru == 8 || ru == 16));
}

int log2i(int k) {
  int l = 0;
  while ((1 << l) < k) ++l;
  return l;
}

// (ia, ja, a, n, k, mode) -> chosen tile
std::map<std::tuple<const void *, const void *, const void *, int, int, int>, int> g_tile;

unsigned spmm_grid() {
  const int per_cu = std::max(1, std::min(16, options().spmm_wg_per_cu));
  return (unsigned)std::max(1, per_cu * device_cus());
}

unsigned vec_grid(size_t elems) {
  return (unsigned)std::max<size_t>(
      1, std::min<size_t>(1024, (elems + 4 * kBlock - 1) / (4 * kBlock)));
}

// Times every rows-per-group choice on Y = A X (three runs each, after a
// warm-up).
int pick_tile(int n, const int *ia, const int *ja, const double *a, int k, const double *X,
              double *Y, int mode = 0, double *part = nullptr, const BlkCtl *ctl = nullptr) {
  auto key = std::make_tuple((const void *)ia, (const void *)ja, (const void *)a, n, k, mode);
  auto it = g_tile.find(key);
  if (it != g_tile.end()) return it->second;
  int best_ru = 0;
  float best = 1e30f;
  const unsigned g = spmm_grid();
  hipEvent_t e0, e1;
  if (!LAP_HIP(hipEventCreate(&e0))) return 0;
  if (!LAP_HIP(hipEventCreate(&e1))) {
    (void)hipEventDestroy(e0);
    return 0;
  }
  for (int ru = 1; ru <= 16; ru *= 2) {
    if (!tile_ok(k, ru)) continue;
    SpmmFn fn = spmm_for(k, ru);
    fn(g, n, ia, ja, a, X, Y, nullptr, part, ctl, mode);
    LAP_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < 3; ++r) fn(g, n, ia, ja, a, X, Y, nullptr, part, ctl, mode);
    LAP_HIP(hipEventRecord(e1, stream()));
    LAP_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    LAP_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      best_ru = ru;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (best_ru) g_tile[key] = best_ru;
  return best_ru;
}

bool block_args(const char *fn, int n, int nnz, const int *ia, const int *ja, const double *a,
                int k, int tile) {
  if (n <= 0 || nnz < 0 || k <= 0 || k > kMaxK || (k & (k - 1))) {
    set_error(LAP_ERR_ARG,
              "%s: need n > 0, nnz >= 0 and k a power of two in [1, %d] (n=%d k=%d)", fn,
              kMaxK, n, k);
    return false;
  }
  if (tile && !tile_ok(k, tile)) {
    set_error(LAP_ERR_ARG, "%s: tile must be 1, 2 or 4 (or 8, 16 with k = 64), not %d", fn, tile);
    return false;
  }
  return check_device(fn, {{ia, "ia", sizeof(int) * ((size_t)n + 1)},
                           {ja, "ja", sizeof(int) * (size_t)nnz},
                           {a, "a", sizeof(double) * (size_t)nnz}});
}

}  // namespace

void release_block_plans() { g_tile.clear(); }
void forget_block_plans(const void *p) {
  for (auto it = g_tile.begin(); it != g_tile.end();) {
    if (std::get<0>(it->first) == p || std::get<1>(it->first) == p ||
        std::get<2>(it->first) == p)
      it = g_tile.erase(it);
    else
      ++it;
  }
}

}  // namespace lap

using namespace lap;

extern "C" int lap_csr_matmat(int n, int nnz, const int *ia, const int *ja, const double *a, int k,
                              const double *X, double *Y, int tile) {
  if (!block_args("csr_matmat", n, nnz, ia, ja, a, k, tile)) return lap_last_status();
  const size_t nk = sizeof(double) * (size_t)n * (size_t)k;
  if (!check_device("csr_matmat", {{X, "X", nk}, {Y, "Y", nk}})) return lap_last_status();
  if (!spmv_plan(n, nnz, ia, ja)) return lap_last_status();  // validates ia / ja
  if (!tile) tile = pick_tile(n, ia, ja, a, k, X, Y);
  if (!tile) return LAP_ERR_HIP;
  spmm_for(k, tile)(spmm_grid(), n, ia, ja, a, X, Y, nullptr, nullptr, nullptr, 0);
  if (!LAP_HIP(hipGetLastError())) return LAP_ERR_HIP;
  sync();
  return LAP_OK;
}

extern "C" int lap_pcg_block(int n, int nnz, const int *ia, const int *ja, const double *a,
                             const double *d_r, int k, double *X, const double *Bm, double tol,
                             int maxit, int tile, double *h_hist, int *iters, int *flags,
                             double *res0, double *res, double *gpu_ms) {
  if (!block_args("pcg_block", n, nnz, ia, ja, a, k, tile)) return lap_last_status();
  if (maxit < 0) {
    set_error(LAP_ERR_ARG, "pcg_block: maxit %d < 0", maxit);
    return LAP_ERR_ARG;
  }
  const size_t nk = sizeof(double) * (size_t)n * (size_t)k;
  if (!check_device("pcg_block", {{d_r, "d_r", sizeof(double) * (size_t)n},
                                  {X, "X", nk},
                                  {Bm, "B", nk}}))
    return lap_last_status();
  if (!spmv_plan(n, nnz, ia, ja)) return lap_last_status();  // validates ia / ja
  double *R = scratch_vec("blk_r", (size_t)n * k), *P = scratch_vec("blk_p", (size_t)n * k),
         *Q = scratch_vec("blk_q", (size_t)n * k);
  BlkCtl *ctl =
      reinterpret_cast<BlkCtl *>(scratch_vec("blk_ctl", sizeof(BlkCtl) / sizeof(double) + 1));
  const unsigned g = spmm_grid(), gv = vec_grid((size_t)n * k);
  double *part = scratch_vec("blk_part", (size_t)2 * kMaxK * std::max(g, gv));
  const size_t hist_len = ((size_t)maxit + 2) * (size_t)k;
  if (h_hist && hist_len > (size_t)INT_MAX) {
    set_error(LAP_ERR_ARG, "pcg_block: history of (maxit+2)*k entries is too long");
    return LAP_ERR_ARG;
  }
  double *hist = h_hist ? scratch_vec("blk_hist", hist_len) : nullptr;
  if (!R || !P || !Q || !ctl || !part || (h_hist && !hist)) return lap_last_status();
  // rows past a column's own stop stay 0
  if (hist) launch_fill((int)hist_len, 0.0, hist);
  if (!tile) {  // timed in the solve's own mode (q = A p with the p.q epilogue)
    const int live = 0;
    LAP_HIP(hipMemcpyAsync(&ctl->all_done, &live, sizeof(int), hipMemcpyHostToDevice, stream()));
    tile = pick_tile(n, ia, ja, a, k, Bm, Q, 1, part, ctl);
  }
  if (!tile) return LAP_ERR_HIP;
  SpmmFn spmm = spmm_for(k, tile);
  const int lk = log2i(k);

  hipEvent_t ev0, ev1;
  if (!LAP_HIP(hipEventCreate(&ev0))) return LAP_ERR_HIP;
  if (!LAP_HIP(hipEventCreate(&ev1))) {
    (void)hipEventDestroy(ev0);
    return LAP_ERR_HIP;
  }
  LAP_HIP(hipEventRecord(ev0, stream()));
  spmm(g, n, ia, ja, a, X, R, Bm, part, ctl, 2);
  k_blk_res0<<<1, kRedThreads, 0, stream()>>>(part, (int)g, k, tol, maxit, ctl, hist);
  k_blk_start<kBlock><<<gv, kBlock, 0, stream()>>>(n, k, lk, R, d_r, P, part);
  k_blk_rho<<<1, kRedThreads, 0, stream()>>>(part, (int)gv, k, ctl);
  auto iteration = [&]() {
    spmm(g, n, ia, ja, a, P, Q, nullptr, part, ctl, 1);
    k_blk_alpha<<<1, kRedThreads, 0, stream()>>>(part, (int)g, k, ctl);
    k_blk_update<kBlock><<<gv, kBlock, 0, stream()>>>(n, k, lk, R, Q, d_r, ctl, part);
    k_blk_conv<<<1, kRedThreads, 0, stream()>>>(part, (int)gv, k, hist, ctl);
    k_blk_dir<kBlock, false><<<gv, kBlock, 0, stream()>>>(n, k, lk, X, R, d_r, P, ctl);
  };
  // iter <= maxit + 1 for every column, so G * max_batches iterations
  // always reach all_done
  const int G = std::max(1, std::min(options().graph_iters, maxit + 2));
  const long max_batches = (long)(maxit + 1) / G + 3;
  int status = LAP_OK;
  hipGraphExec_t gexec = nullptr;
  if (options().use_graph) {
    hipGraph_t graph = nullptr;
    bool ok = LAP_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
    if (ok) {
      for (int s = 0; s < G; ++s) iteration();
      ok = LAP_HIP(hipStreamEndCapture(stream(), &graph));
    }
    if (ok) ok = LAP_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    if (graph) (void)hipGraphDestroy(graph);
    if (!ok) status = LAP_ERR_HIP;
  }
  int *flag_h = pinned_flags();
  if (!flag_h) status = LAP_ERR_HIP;
  if (status == LAP_OK) {
    bool finished = false;
    for (long b = 0; b < max_batches && !finished; ++b) {
      if (gexec) {
        if (!LAP_HIP(hipGraphLaunch(gexec, stream()))) break;
      } else {
        for (int s = 0; s < G; ++s) iteration();
      }
      LAP_HIP(hipMemcpyAsync(flag_h, &ctl->all_done, sizeof(int), hipMemcpyDeviceToHost,
                             stream()));
      sync();
      finished = flag_h[0] != 0;
    }
    if (!finished) {
      set_error(LAP_ERR_NUMERIC, "pcg_block: stop flags never set");
      status = LAP_ERR_NUMERIC;
    }
    k_blk_dir<kBlock, true><<<gv, kBlock, 0, stream()>>>(n, k, lk, X, R, d_r, P, ctl);
  }
  LAP_HIP(hipEventRecord(ev1, stream()));
  sync();
  float ms = 0.f;
  LAP_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (gpu_ms) *gpu_ms = ms;
  BlkCtl h{};
  LAP_HIP(hipMemcpy(&h, ctl, sizeof h, hipMemcpyDeviceToHost));
  for (int j = 0; j < k; ++j) {
    if (iters) iters[j] = h.iters[j];
    if (flags) flags[j] = h.flag[j];
    if (res0) res0[j] = h.res0[j];
    if (res) res[j] = h.res[j];
  }
  if (h_hist && hist)
    LAP_HIP(hipMemcpy(h_hist, hist, sizeof(double) * ((size_t)h.iter + 1) * k,
                      hipMemcpyDeviceToHost));
  if (status == LAP_OK && lap_last_status() == LAP_ERR_HIP) status = LAP_ERR_HIP;
  return status;
}
