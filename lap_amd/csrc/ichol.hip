// ichol.hip -- IC(0) on the device (initialize_ichol, hip_ichol).
//
// Factorisation: up-looking, one lane per row, rows in the level order of
// L's pattern (a row depends on the rows of its off-diagonal columns).  For
// row j and each column c of its pattern, ascending:
//     t = A(j,c) - sum_{i<c, i in pat(j) and pat(c), ascending} L(c,i)*L(j,i)
//     L(j,c) = t / L(c,c)  (c < j),   L(j,j) = sqrt(t).
// Each entry receives exactly the subtractions, in the same order, that the
// reference's right-looking CPU sweep applies to it (simple_blas.c:148-168:
// U(r,c) -= U(i,r)*U(i,c) for ascending i, then the division by the pivot
// when row r is reached), so with contraction off the factor is bitwise the
// reference's.  The factor is written into the lower triangle of the
// caller's array (rocSPARSE csric0 semantics, hip_blas.cpp:175-317) and kept
// as L (diagonal last) and U = L^T (diagonal first) for row-oriented
// level-scheduled solves in simple_ichol's order (simple_blas.c:184-223).
#include <algorithm>
#include <cmath>

#include <tuple>

#include "lap_core.hpp"

namespace lap {

namespace {

__global__ __launch_bounds__(kBlock) void k_gather(int cnt, const int *__restrict__ pos,
                                                   const double *__restrict__ src,
                                                   double *__restrict__ dst) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < cnt) dst[t] = src[pos[t]];
}

__global__ __launch_bounds__(kBlock) void k_scatter(int cnt, const int *__restrict__ pos,
                                                    const double *__restrict__ src,
                                                    double *__restrict__ dst) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < cnt) dst[pos[t]] = src[t];
}

// One row j of the factor by one lane (entries ascending; each merges row
// j's earlier entries with row c's off-diagonal entries).
// bad[0]: count of non-positive / non-finite pivots, bad[1]: first such row
__device__ void ic0_row(int j, const int *__restrict__ lia, const int *__restrict__ lja,
                        double *__restrict__ lval, int *__restrict__ bad) {
  const int p0 = lia[j], p1 = lia[j + 1];
  for (int p = p0; p < p1; ++p) {
    const int c = lja[p];
    double acc = lval[p];
    // rows j (entries before p) and c (entries before its diagonal) merge
    int q = p0, s = (c == j) ? p0 : lia[c];
    const int qe = p, se = (c == j) ? p : lia[c + 1] - 1;
    while (q < qe && s < se) {
      const int cq = lja[q], cs = lja[s];
      if (cq == cs) {
        acc -= lval[s] * lval[q];
        ++q;
        ++s;
      } else if (cq < cs) {
        ++q;
      } else {
        ++s;
      }
    }
    if (c == j) {
      if (!(acc > 0.0) || !isfinite(acc)) {
        if (atomicAdd(&bad[0], 1) == 0) bad[1] = j;
      }
      lval[p] = __builtin_sqrt(acc);
    } else {
      lval[p] = acc / lval[lia[c + 1] - 1];
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_ic0_level(int cnt, const int *__restrict__ rows,
                                                      const int *__restrict__ lia,
                                                      const int *__restrict__ lja,
                                                      double *__restrict__ lval,
                                                      int *__restrict__ bad) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= cnt) return;
  ic0_row(rows[t], lia, lja, lval, bad);
}

// Rows of more than kTriLongRow entries: a workgroup per row.  Row j's
// columns and values sit in LDS; its entries are taken in order, and for
// entry (j, c) every thread takes entries of row c (off-diagonal), finds
// their column among row j's earlier entries by binary search and writes
// the product L(c,i) * L(j,i) -- or +0.0 where row j has no such column --
// into LDS at the entry's position; lane 0 then subtracts them in row c's
// (ascending) order.  The matched products come in the merge's order and
// x - (+0.0) == x bit for bit, so every entry is exactly ic0_row's.  Rows
// longer than kIcRow entries are left to one lane (ic0_row).
constexpr int kIcRow = 2048;

__global__ __launch_bounds__(kBlock) void k_ic0_long(const int *__restrict__ rows,
                                                     const int *__restrict__ lia,
                                                     const int *__restrict__ lja,
                                                     double *__restrict__ lval,
                                                     int *__restrict__ bad) {
  __shared__ int cj[kIcRow];
  __shared__ double vj[kIcRow];
  __shared__ double prod[kIcRow];
  const int j = __builtin_amdgcn_readfirstlane(rows[blockIdx.x]);
  const int p0 = __builtin_amdgcn_readfirstlane(lia[j]);
  const int len = __builtin_amdgcn_readfirstlane(lia[j + 1]) - p0;
  if (len > kIcRow) {
    if (threadIdx.x == 0) ic0_row(j, lia, lja, lval, bad);
    return;
  }
  for (int k = threadIdx.x; k < len; k += kBlock) {
    cj[k] = lja[p0 + k];
    vj[k] = lval[p0 + k];  // A's values; each is replaced once computed
  }
  __syncthreads();
  for (int pk = 0; pk < len; ++pk) {
    const int c = __builtin_amdgcn_readfirstlane(cj[pk]);
    double acc = vj[pk];
    if (c == j) {  // the diagonal: minus the squares of the row's entries
      for (int k = threadIdx.x; k < pk; k += kBlock) prod[k] = vj[k] * vj[k];
      __syncthreads();
      if (threadIdx.x == 0) {
        acc = lds_seq_sum<true>(acc, prod, 0, pk);
        if (!(acc > 0.0) || !isfinite(acc)) {
          if (atomicAdd(&bad[0], 1) == 0) bad[1] = j;
        }
        const double v = __builtin_sqrt(acc);
        vj[pk] = v;
        lval[p0 + pk] = v;
      }
    } else {
      const int s0 = __builtin_amdgcn_readfirstlane(lia[c]);
      const int lenc = __builtin_amdgcn_readfirstlane(lia[c + 1]) - 1 - s0;
      for (int base = 0; base < lenc; base += kIcRow) {
        const int m = min(kIcRow, lenc - base);
        for (int k = threadIdx.x; k < m; k += kBlock) {
          const int s = s0 + base + k;
          const int col = lja[s];
          int lo = 0, hi = pk;  // row j's entries before (j, c), ascending
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (cj[mid] < col)
              lo = mid + 1;
            else
              hi = mid;
          }
          prod[k] = (lo < pk && cj[lo] == col) ? lval[s] * vj[lo] : 0.0;
        }
        __syncthreads();
        if (threadIdx.x == 0) acc = lds_seq_sum<true>(acc, prod, 0, m);
        __syncthreads();
      }
      if (threadIdx.x == 0) {
        const double v = acc / lval[s0 + lenc];
        vj[pk] = v;
        lval[p0 + pk] = v;
      }
    }
    __syncthreads();
  }
}

struct IcEntry {
  std::unique_ptr<IcPlan> plan;
  std::unique_ptr<TriPlan> lplan, uplan;
  const int *ia = nullptr, *ja = nullptr;
};
std::map<const double *, IcEntry> g_ic;

}  // namespace

void release_ic_plans() { g_ic.clear(); }
void forget_ic_plans(const void *p) {
  for (auto it = g_ic.begin(); it != g_ic.end();) {
    if (it->first == p || it->second.ia == p || it->second.ja == p)
      it = g_ic.erase(it);
    else
      ++it;
  }
}

const IcPlan *ic_plan_find(const double *vals) {
  auto it = g_ic.find(vals);
  return it == g_ic.end() ? nullptr : it->second.plan.get();
}

// Level analyses of a lower and an upper pattern held on the host (shares
// the solve kernels' TriPlan layout), the two at once.
std::pair<std::unique_ptr<TriPlan>, std::unique_ptr<TriPlan>> tri_analyse_pair(
    int n, const std::vector<int> &lia, const std::vector<int> &lja,
    const std::vector<int> &uia, const std::vector<int> &uja);

const IcPlan *ic_factor(int n, int nnzA, const int *ia, const int *ja, double *a) {
  if (n < 0 || nnzA < 0) {
    set_error(LAP_ERR_ARG, "initialize_ichol: negative size");
    return nullptr;
  }
  std::vector<int> h_ia = dev_download(ia, (size_t)n + 1);
  const int total = n ? h_ia[n] : 0;
  if (n && (h_ia[0] != 0 || total > nnzA)) {
    set_error(LAP_ERR_STRUCT, "initialize_ichol: row pointers inconsistent with nnzA=%d", nnzA);
    return nullptr;
  }
  std::vector<int> h_ja = dev_download(ja, (size_t)total);

  // L pattern = lower triangle of A incl. diagonal; lpos maps into A
  std::vector<int> lia((size_t)n + 1, 0), lja, lpos;
  lja.reserve((size_t)total / 2 + n);
  lpos.reserve((size_t)total / 2 + n);
  for (int i = 0; i < n; ++i) {
    int last = -1;
    bool diag = false;
    for (int e = h_ia[i]; e < h_ia[i + 1]; ++e) {
      const int c = h_ja[e];
      if (c < 0 || c >= n || c <= last) {
        set_error(LAP_ERR_STRUCT, "initialize_ichol: row %d not sorted / column out of range", i);
        return nullptr;
      }
      last = c;
      if (c <= i) {
        lja.push_back(c);
        lpos.push_back(e);
        diag = diag || (c == i);
      }
    }
    if (!diag) {
      set_error(LAP_ERR_STRUCT, "A has structural zero at A(%d,%d)", i, i);
      return nullptr;
    }
    lia[i + 1] = (int)lja.size();
  }
  const int nnzL = (int)lja.size();
  // U = L^T pattern and the permutation U entry -> L entry
  std::vector<int> uia((size_t)n + 1, 0), uja((size_t)nnzL), tperm((size_t)nnzL);
  for (int k = 0; k < nnzL; ++k) uia[lja[k] + 1]++;
  for (int i = 0; i < n; ++i) uia[i + 1] += uia[i];
  {
    std::vector<int> fill(uia.begin(), uia.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int k = lia[i]; k < lia[i + 1]; ++k) {
        const int slot = fill[lja[k]]++;
        uja[slot] = i;
        tperm[slot] = k;
      }
  }

  IcEntry ent;
  ent.plan = std::make_unique<IcPlan>();
  IcPlan &P = *ent.plan;
  P.n = n;
  P.nnzL = nnzL;
  P.lia = dev_upload(lia);
  P.lja = dev_upload(lja);
  P.uia = dev_upload(uia);
  P.uja = dev_upload(uja);
  P.lval = dev_alloc(sizeof(double) * (size_t)std::max(nnzL, 1));
  P.uval = dev_alloc(sizeof(double) * (size_t)std::max(nnzL, 1));
  auto d_lpos = dev_upload(lpos);
  auto d_tperm = dev_upload(tperm);
  std::tie(ent.lplan, ent.uplan) = tri_analyse_pair(n, lia, lja, uia, uja);
  if (!ent.lplan || !ent.uplan) return nullptr;
  P.lplan = ent.lplan.get();
  P.uplan = ent.uplan.get();

  auto bad = dev_alloc(2 * sizeof(int));
  LAP_HIP(hipMemsetAsync(bad->p, 0, 2 * sizeof(int), stream()));
  const unsigned gL = (unsigned)((nnzL + kBlock - 1) / kBlock);
  if (nnzL) k_gather<<<gL, kBlock, 0, stream()>>>(nnzL, d_lpos->as<int>(), a, P.lval->as<double>());
  const int *order = P.lplan->order->as<int>();
  for (int l = 0; l < P.lplan->nlevels; ++l) {
    // the level's short rows (one lane each), then its long rows (one
    // workgroup each); rows of one level never depend on each other
    const int l0 = P.lplan->level_ptr[l], ll = P.lplan->level_long[l];
    const int cnt = ll - l0, nlong = P.lplan->level_ptr[l + 1] - ll;
    if (cnt > 0)
      k_ic0_level<<<(cnt + kBlock - 1) / kBlock, kBlock, 0, stream()>>>(
          cnt, order + l0, P.lia->as<int>(), P.lja->as<int>(), P.lval->as<double>(),
          bad->as<int>());
    if (nlong > 0)
      k_ic0_long<<<nlong, kBlock, 0, stream()>>>(order + ll, P.lia->as<int>(), P.lja->as<int>(),
                                                 P.lval->as<double>(), bad->as<int>());
  }
  if (nnzL) {
    k_scatter<<<gL, kBlock, 0, stream()>>>(nnzL, d_lpos->as<int>(), P.lval->as<double>(), a);
    k_gather<<<gL, kBlock, 0, stream()>>>(nnzL, d_tperm->as<int>(), P.lval->as<double>(),
                                          P.uval->as<double>());
  }
  tri_ell_build(*P.lplan, P.lja->as<int>(), P.lval->as<double>(), P.lell);
  tri_ell_build(*P.uplan, P.uja->as<int>(), P.uval->as<double>(), P.uell);
  std::vector<int> h_bad = dev_download(bad->as<int>(), 2);
  if (h_bad[0]) {
    P.ok = false;
    set_error(LAP_ERR_NUMERIC, "L has structural and/or numerical zero at L(%d,%d)", h_bad[1],
              h_bad[1]);
  }
  ent.ia = ia;
  ent.ja = ja;
  const IcPlan *raw = ent.plan.get();
  g_ic[a] = std::move(ent);
  return raw;
}

void ic_apply_launch(const IcPlan &p, const double *x, double *y, double *aux,
                     const int *done) {
  tri_solve_launch(*p.lplan, p.lia->as<int>(), p.lja->as<int>(), p.lval->as<double>(), nullptr,
                   x, aux, done, &p.lell);
  tri_solve_launch(*p.uplan, p.uia->as<int>(), p.uja->as<int>(), p.uval->as<double>(), nullptr,
                   aux, y, done, &p.uell);
}

}  // namespace lap

using namespace lap;

extern "C" {

void initialize_ichol(const int n, const int nnzA, int *ia, int *ja, double *a) {
  if (n <= 0 || !check_device("initialize_ichol", {{ia, "ia"}, {ja, "ja"}, {a, "a"}})) return;
  ic_factor(n, nnzA, ia, ja, a);
  sync();
}

void hip_ichol(const int *ia, const int *ja, double *a, const int nnzA, lap_pdata *prec_data,
               double *x, double *y) {
  (void)ia;
  (void)ja;
  (void)a;
  (void)nnzA;
  const IcPlan *p = ic_plan_find(prec_data->ichol_vals);
  if (!p) {
    set_error(LAP_ERR_ARG, "hip_ichol: no factor for ichol_vals=%p (call initialize_ichol)",
              (void *)prec_data->ichol_vals);
    return;
  }
  if (!check_device("hip_ichol", {{x, "x"}, {y, "y"}})) return;
  double *aux = is_device_ptr(prec_data->aux_vec1) ? prec_data->aux_vec1
                                                   : scratch_vec("ichol_aux", p->n);
  ic_apply_launch(*p, x, y, aux, nullptr);
  sync();
  if (options().sptrsv_mode != 0 && tri_error()) {
    set_error(LAP_ERR_NUMERIC, "hip_ichol: a triangular-solve dependency wait ran out");
    tri_clear_error();
  }
}

void ichol(const int *ia, const int *ja, double *a, const int nnzA, lap_pdata *prec_data,
           double *x, double *y) {
  hip_ichol(ia, ja, a, nnzA, prec_data, x, y);
}

}  // extern "C"
