// adjoint.hip -- weighted graph Laplacians assembled on the device from
// edge weights, and the adjoint gradient of a PCG solve with respect to the
// edge weights and the right-hand side.  No reference counterpart: the
// reference builds an unweighted Laplacian once on the host
// (io_utils.c:80-130, :364-420) and has no derivative of the solve.
//
// System: L(w) x = b with L(w) = sum_e w_e (u_e - v_e)(u_e - v_e)^T + shift*I
// over undirected edges e = (s_e, d_e).  For a loss l(x):
//   adjoint   L(w) lambda = dl/dx            (L is symmetric)
//   gradients dl/db = lambda,  dl/dw_e = -(lambda_s - lambda_d)(x_s - x_d)
// The adjoint solve is one more PCG with the forward solve's preconditioner
// (lap_pcg), followed by one fused pass over the edges.
//
// CSR pattern (lap_graph_pattern): rows sorted by column, the diagonal
// always present, epos[2e] / epos[2e+1] the positions of (s_e, d_e) and
// (d_e, s_e), dpos[i] the position of (i, i).  Values (lap_graph_values):
// off-diagonals -w_e; the diagonal is the sum of the row's weights taken in
// CSR order, plus shift.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "lap_core.hpp"

namespace lap {

namespace {

// Positions outside [0, nnz) (a map not made by lap_graph_pattern) are
// skipped rather than written.
__global__ __launch_bounds__(kBlock) void k_scatter_weights(int m, unsigned nnz,
                                                            const int *__restrict__ epos,
                                                            const double *__restrict__ w,
                                                            double *__restrict__ a) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m) return;
  const double v = -w[e];
  const unsigned p0 = (unsigned)epos[2 * e], p1 = (unsigned)epos[2 * e + 1];
  if (p0 < nnz) a[p0] = v;
  if (p1 < nnz) a[p1] = v;
}

// diagonal = (sum of the row's weights in CSR order) + shift; d_r = 1/diag
// (0 where diag == 0, as vector_reciprocal)
__global__ __launch_bounds__(kBlock) void k_graph_diag(int n, unsigned nnz,
                                                       const int *__restrict__ ia,
                                                       const int *__restrict__ ja,
                                                       const int *__restrict__ dpos,
                                                       double shift, double *__restrict__ a,
                                                       double *__restrict__ d_r) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  const int e1 = ia[i + 1];
  for (int e = ia[i]; e < e1; ++e)
    if (ja[e] != i) s += -a[e];
  const double dg = s + shift;
  const unsigned pd = (unsigned)dpos[i];
  if (pd < nnz) a[pd] = dg;
  if (d_r) d_r[i] = dg != 0.0 ? 1.0 / dg : 0.0;
}

// g_e = -(lambda_s - lambda_d) * (x_s - x_d); NaN for an endpoint outside [0, n)
__global__ __launch_bounds__(kBlock) void k_edge_grad(int m, unsigned n,
                                                      const int *__restrict__ src,
                                                      const int *__restrict__ dst,
                                                      const double *__restrict__ lam,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ g) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m) return;
  const unsigned s = (unsigned)src[e], d = (unsigned)dst[e];
  g[e] = (s < n && d < n) ? -((lam[s] - lam[d]) * (x[s] - x[d])) : __builtin_nan("");
}

inline unsigned blocks_for(int cnt) { return (unsigned)((cnt + kBlock - 1) / kBlock); }

}  // namespace

}  // namespace lap

using namespace lap;

extern "C" {

int lap_graph_pattern(int n, int m, const int *src, const int *dst, int *ia, int *ja, int *epos,
                      int *dpos) {
  if (n <= 0 || m < 0 || (long long)n + 2LL * m > (long long)INT32_MAX) {
    set_error(LAP_ERR_ARG, "graph_pattern: bad sizes n=%d m=%d", n, m);
    return LAP_ERR_ARG;
  }
  const size_t nnz = (size_t)n + 2 * (size_t)m;
  if (!check_device("graph_pattern", {{src, "src", sizeof(int) * (size_t)m},
                                      {dst, "dst", sizeof(int) * (size_t)m},
                                      {ia, "ia", sizeof(int) * ((size_t)n + 1)},
                                      {ja, "ja", sizeof(int) * nnz},
                                      {epos, "epos", sizeof(int) * 2 * (size_t)m},
                                      {dpos, "dpos", sizeof(int) * (size_t)n}}))
    return LAP_ERR_ARG;
  std::vector<int> s = dev_download(src, (size_t)m), d = dev_download(dst, (size_t)m);
  std::vector<int> cnt((size_t)n + 1, 0);
  for (int e = 0; e < m; ++e) {
    if (s[e] < 0 || s[e] >= n || d[e] < 0 || d[e] >= n || s[e] == d[e]) {
      set_error(LAP_ERR_STRUCT, "graph_pattern: edge %d (%d, %d) is a self loop or out of range",
                e, s[e], d[e]);
      return LAP_ERR_STRUCT;
    }
    cnt[s[e] + 1]++;
    cnt[d[e] + 1]++;
  }
  std::vector<int> h_ia((size_t)n + 1, 0);
  for (int i = 0; i < n; ++i) h_ia[i + 1] = h_ia[i] + cnt[i + 1] + 1;
  // entries per row: (column, tag) with tag = 2e / 2e+1 for edge slots, -1 diagonal
  std::vector<std::pair<int, int>> ent(nnz);
  std::vector<int> fill(h_ia.begin(), h_ia.end() - 1);
  for (int i = 0; i < n; ++i) ent[fill[i]++] = {i, -1};
  for (int e = 0; e < m; ++e) {
    ent[fill[s[e]]++] = {d[e], 2 * e};
    ent[fill[d[e]]++] = {s[e], 2 * e + 1};
  }
  std::vector<int> h_ja(nnz), h_epos(2 * (size_t)m), h_dpos((size_t)n);
  for (int i = 0; i < n; ++i) {
    auto b = ent.begin() + h_ia[i], en = ent.begin() + h_ia[i + 1];
    std::sort(b, en);
    for (int p = h_ia[i]; p < h_ia[i + 1]; ++p) {
      if (p > h_ia[i] && ent[p].first == ent[p - 1].first) {
        set_error(LAP_ERR_STRUCT, "graph_pattern: duplicate edge between %d and %d", i,
                  ent[p].first);
        return LAP_ERR_STRUCT;
      }
      h_ja[p] = ent[p].first;
      if (ent[p].second < 0)
        h_dpos[i] = p;
      else
        h_epos[ent[p].second] = p;
    }
  }
  bool ok = LAP_HIP(hipMemcpy(ia, h_ia.data(), sizeof(int) * h_ia.size(), hipMemcpyHostToDevice));
  ok = ok && LAP_HIP(hipMemcpy(ja, h_ja.data(), sizeof(int) * nnz, hipMemcpyHostToDevice));
  if (m)
    ok = ok && LAP_HIP(hipMemcpy(epos, h_epos.data(), sizeof(int) * h_epos.size(),
                                 hipMemcpyHostToDevice));
  ok = ok && LAP_HIP(hipMemcpy(dpos, h_dpos.data(), sizeof(int) * (size_t)n,
                               hipMemcpyHostToDevice));
  forget_plans(ia);  // a new pattern may reuse addresses of an old one
  forget_plans(ja);
  return ok ? LAP_OK : LAP_ERR_HIP;
}

int lap_graph_values(int n, int m, const int *ia, const int *ja, const int *epos,
                     const int *dpos, const double *w, double shift, double *a, double *d_r) {
  if (n <= 0 || m < 0) {
    set_error(LAP_ERR_ARG, "graph_values: bad sizes n=%d m=%d", n, m);
    return LAP_ERR_ARG;
  }
  const size_t nnz = (size_t)n + 2 * (size_t)m;
  if (!check_device("graph_values", {{ia, "ia", sizeof(int) * ((size_t)n + 1)},
                                     {ja, "ja", sizeof(int) * nnz},
                                     {epos, "epos", sizeof(int) * 2 * (size_t)m},
                                     {dpos, "dpos", sizeof(int) * (size_t)n},
                                     {w, "w", sizeof(double) * (size_t)m},
                                     {a, "a", sizeof(double) * nnz}}))
    return LAP_ERR_ARG;
  if (d_r && !check_device("graph_values", {{d_r, "d_r", sizeof(double) * (size_t)n}}))
    return LAP_ERR_ARG;
  // validates ia (monotone, ending at nnz) and ja (in range) once per pattern
  if (!spmv_plan(n, (int)nnz, ia, ja)) return lap_last_status();
  if (m) k_scatter_weights<<<blocks_for(m), kBlock, 0, stream()>>>(m, (unsigned)nnz, epos, w, a);
  k_graph_diag<<<blocks_for(n), kBlock, 0, stream()>>>(n, (unsigned)nnz, ia, ja, dpos, shift, a,
                                                       d_r);
  if (!LAP_HIP(hipGetLastError())) return LAP_ERR_HIP;
  sync();
  return LAP_OK;
}

int lap_edge_grad(int n, int m, const int *src, const int *dst, const double *lam,
                  const double *x, double *grad_w) {
  if (n <= 0 || m < 0) {
    set_error(LAP_ERR_ARG, "edge_grad: bad sizes n=%d m=%d", n, m);
    return LAP_ERR_ARG;
  }
  if (!m) return LAP_OK;
  if (!check_device("edge_grad", {{src, "src", sizeof(int) * (size_t)m},
                                  {dst, "dst", sizeof(int) * (size_t)m},
                                  {lam, "lambda", sizeof(double) * (size_t)n},
                                  {x, "x", sizeof(double) * (size_t)n},
                                  {grad_w, "grad_w", sizeof(double) * (size_t)m}}))
    return LAP_ERR_ARG;
  k_edge_grad<<<blocks_for(m), kBlock, 0, stream()>>>(m, (unsigned)n, src, dst, lam, x, grad_w);
  if (!LAP_HIP(hipGetLastError())) return LAP_ERR_HIP;
  sync();
  return LAP_OK;
}

int lap_pcg_adjoint(int n, int nnz, const int *ia, const int *ja, const double *a,
                    lap_pdata *prec_data, const double *grad_x, double *lam, double tol,
                    int maxit, int m, const int *src, const int *dst, const double *x,
                    double *grad_w, double *history, lap_solve_info *info) {
  if (n <= 0) {
    set_error(LAP_ERR_ARG, "pcg_adjoint: n=%d", n);
    return LAP_ERR_ARG;
  }
  if (!check_device("pcg_adjoint", {{lam, "lambda", sizeof(double) * (size_t)n}}))
    return LAP_ERR_ARG;
  launch_fill(n, 0.0, lam);  // lambda_0 = 0
  lap_solve_info local{};
  lap_solve_info *out = info ? info : &local;
  const int st = lap_pcg(n, nnz, ia, ja, a, lam, grad_x, tol, prec_data, maxit, history, out);
  if (st != LAP_OK) return st;
  if (out->flag != 0) {
    set_error(LAP_ERR_NUMERIC, "pcg_adjoint: adjoint solve stopped with flag %d", out->flag);
    return LAP_ERR_NUMERIC;
  }
  return grad_w ? lap_edge_grad(n, m, src, dst, lam, x, grad_w) : LAP_OK;
}

}  // extern "C"
