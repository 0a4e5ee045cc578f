// sweep.hip -- one launch per sweep of the SpMV-only preconditioners
// it_jacobi (it_jacobi.c:4-32), GS_it (GS.c:45-92) and GS_it2 (GS.c:95-159).
//
// The reference runs every sweep as two or three SpMVs plus copies, axpys
// and an element-wise product, each a pass over n-vectors.  Every sweep is
// one of two forms, and each is one kernel here:
//
//   single  y_i = base_i - sum_T a_ij x_j      (sum in CSR order, starting
//           from base_i: csr_matvec with alpha = -1, beta = 1 on a copy of
//           base), then stored as y_i, y_i * d_r_i, or both, or added into
//           w (w_i + y_i * d_r_i);
//   pair    y_i = (base_i - (sum_U a_ij x_j + sum_L a_ij x_j)) * d_r_i
//           (two SpMVs from zero, their sum, the difference, the scaling).
//
// Both use the row-block SpMV body: nonzeros are read coalesced into LDS
// products, then one lane per row sums its products in CSR order, so every
// value equals the unfused sequence bit for bit (rows longer than the LDS
// tile use the same workgroup tree as the plain SpMV).  A pair sweep walks
// row blocks of L and U together (PairPlan: at most kSpmvTile nonzeros of
// L and U per block).
#include <algorithm>

#include "lap_core.hpp"
#include "spmv_core.hpp"

namespace lap {

namespace {

// single-form store modes
enum SweepStore : int {
  kStoreScaled = 0,  // out = y * d_r
  kStorePlain = 1,   // out = y
  kStoreBoth = 2,    // out = y, out2 = y * d_r
  kStoreAccum = 3,   // out2 = (first ? 0 : out2) + y * d_r
};

template <int MODE>
struct SweepEpi {
  const double *base, *d_r;
  double *out, *out2;
  int first;
  __device__ double init(int row) const { return base[row]; }
  __device__ void store(int row, double v) const {
    if constexpr (MODE == kStoreScaled) {
      out[row] = v * d_r[row];
    } else if constexpr (MODE == kStorePlain) {
      out[row] = v;
    } else if constexpr (MODE == kStoreBoth) {
      out[row] = v;
      out2[row] = v * d_r[row];
    } else {
      const double s = v * d_r[row];
      out2[row] = (first ? 0.0 : out2[row]) + s;
    }
  }
};

template <int BS, int TILE, int MODE>
__global__ __launch_bounds__(BS) void k_sweep_one(int nblk, const int *__restrict__ rowblk,
                                                  const int *__restrict__ ia,
                                                  const int *__restrict__ ja,
                                                  const double *__restrict__ a,
                                                  const double *__restrict__ x,
                                                  const double *base, const double *d_r,
                                                  double *out, double *out2, int first, int xcd,
                                                  const int *done) {
  if (done && *done) return;
  __shared__ double prod[TILE];
  __shared__ double red[BS / kWave];
  SweepEpi<MODE> epi{base, d_r, out, out2, first};
  spmv_walk<BS, TILE>(block_walk(nblk, xcd), rowblk, ia, ja, a, x, -1.0, prod, red, epi);
}

// Sum of a[e] * x[ja[e]] over [e0, e1) from 0.0, as the plain SpMV sums one
// row: in CSR order through LDS when it fits the tile, else the workgroup
// tree of spmv_long_row.  All threads call it; the result is returned to
// all of them.
template <int BS, int TILE>
__device__ double row_sum_any(int e0, int e1, const int *__restrict__ ja,
                              const double *__restrict__ a, const double *__restrict__ x,
                              double *prod, double *red) {
  const int cnt = e1 - e0;
  double r;
  if (cnt > TILE) {
    double s = 0.0;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += BS) s += (1.0 * a[e]) * x[ja[e]];
    const double t = block_sum<BS>(s, red);
    if (threadIdx.x == 0) red[0] = 0.0 + t;
  } else {
    for (int k = threadIdx.x; k < cnt; k += BS) prod[k] = (1.0 * a[e0 + k]) * x[ja[e0 + k]];
    __syncthreads();
    if (threadIdx.x == 0) {
      red[0] = lds_seq_sum(0.0, prod, 0, cnt);
    }
  }
  __syncthreads();
  r = red[0];
  __syncthreads();
  return r;
}

template <int BS, int TILE>
__global__ __launch_bounds__(BS) void k_sweep_pair(
    int nblk, const int *__restrict__ rowblk, const int *__restrict__ lia,
    const int *__restrict__ lja, const double *__restrict__ la, const int *__restrict__ uia,
    const int *__restrict__ uja, const double *__restrict__ ua, const double *__restrict__ x,
    const double *base, const double *d_r, double *out, int xcd, const int *done) {
  if (done && *done) return;
  __shared__ double prod[TILE];
  __shared__ double red[BS / kWave];
  const BlockWalk w = block_walk(nblk, xcd);
  for (int s = w.begin; s < w.end; s += w.step) {
    const int b = w.block(s);
    if (b >= nblk) continue;
    const int r0 = rowblk[b], r1 = rowblk[b + 1];
    const int l0 = lia[r0], l1 = lia[r1], u0 = uia[r0], u1 = uia[r1];
    const int cl = l1 - l0, cu = u1 - u0;
    if (cl + cu > TILE) {  // a row of its own (the plan guarantees r1 == r0 + 1)
      const double sl = row_sum_any<BS, TILE>(l0, l1, lja, la, x, prod, red);
      const double su = row_sum_any<BS, TILE>(u0, u1, uja, ua, x, prod, red);
      if (threadIdx.x == 0) out[r0] = (base[r0] - (su + sl)) * d_r[r0];
      continue;
    }
#pragma unroll 4
    for (int k = threadIdx.x; k < cl; k += BS) prod[k] = (1.0 * la[l0 + k]) * x[lja[l0 + k]];
#pragma unroll 4
    for (int k = threadIdx.x; k < cu; k += BS)
      prod[cl + k] = (1.0 * ua[u0 + k]) * x[uja[u0 + k]];
    __syncthreads();
    for (int row = r0 + (int)threadIdx.x; row < r1; row += BS) {
      const int a0 = lia[row] - l0, a1 = lia[row + 1] - l0;
      const double sl = lds_seq_sum(0.0, prod, a0, a1);
      const int c0 = uia[row] - u0 + cl, c1 = uia[row + 1] - u0 + cl;
      const double su = lds_seq_sum(0.0, prod, c0, c1);
      out[row] = (base[row] - (su + sl)) * d_r[row];
    }
    __syncthreads();
  }
}

using PairKey = std::tuple<const int *, const int *, int, int, int, int>;
std::map<PairKey, std::unique_ptr<PairPlan>> g_pair;

}  // namespace

const PairPlan *pair_plan(int n, int nl, const int *lia, int nu, const int *uia) {
  const int R = options().spmv_rows;
  PairKey key{lia, uia, n, nl, nu, R};
  auto it = g_pair.find(key);
  if (it != g_pair.end()) return it->second.get();
  const std::vector<int> hl = dev_download(lia, (size_t)n + 1), hu = dev_download(uia, (size_t)n + 1);
  std::vector<int> starts{0};
  int r = 0;
  while (r < n) {
    const int len = (hl[r + 1] - hl[r]) + (hu[r + 1] - hu[r]);
    if (len > kSpmvTile) {
      starts.push_back(++r);
      continue;
    }
    int r1 = r + 1;
    while (r1 < n && r1 - r < R &&
           (hl[r1 + 1] - hl[r]) + (hu[r1 + 1] - hu[r]) <= kSpmvTile)
      ++r1;
    starts.push_back(r1);
    r = r1;
  }
  auto plan = std::make_unique<PairPlan>();
  plan->n = n;
  plan->nblocks = (int)starts.size() - 1;
  plan->rowblk = dev_upload(starts);
  PairPlan *raw = plan.get();
  g_pair[key] = std::move(plan);
  return raw;
}

void forget_pair_plans(const void *p) {
  for (auto it = g_pair.begin(); it != g_pair.end();) {
    if (std::get<0>(it->first) == p || std::get<1>(it->first) == p)
      it = g_pair.erase(it);
    else
      ++it;
  }
}
void release_pair_plans() { g_pair.clear(); }

void sweep_one_launch(const SpmvPlan &p, const int *ia, const int *ja, const double *a,
                      const double *x, const double *base, const double *d_r, double *out,
                      double *out2, int mode, bool first, const int *done) {
  if (p.nblocks == 0) return;
  const unsigned g = spmv_grid(p);
  const int xcd = options().spmv_walk;
  const int *rb = p.rowblk->as<int>();
#define LAP_SWEEP_CASE(M)                                                                   \
  case M:                                                                                   \
    k_sweep_one<kBlock, kSpmvTile, M><<<g, kBlock, 0, stream()>>>(                          \
        p.nblocks, rb, ia, ja, a, x, base, d_r, out, out2, first ? 1 : 0, xcd, done);        \
    break;
  switch (mode) {
    LAP_SWEEP_CASE(kStoreScaled)
    LAP_SWEEP_CASE(kStorePlain)
    LAP_SWEEP_CASE(kStoreBoth)
    LAP_SWEEP_CASE(kStoreAccum)
    default:
      break;
  }
#undef LAP_SWEEP_CASE
}

void sweep_pair_launch(const PairPlan &p, const int *lia, const int *lja, const double *la,
                       const int *uia, const int *uja, const double *ua, const double *x,
                       const double *base, const double *d_r, double *out, const int *done) {
  if (p.nblocks == 0) return;
  const unsigned g = (unsigned)std::max(1, std::min(p.nblocks, 8 * device_cus()));
  k_sweep_pair<kBlock, kSpmvTile><<<g, kBlock, 0, stream()>>>(
      p.nblocks, p.rowblk->as<int>(), lia, lja, la, uia, uja, ua, x, base, d_r, out,
      options().spmv_walk, done);
}

}  // namespace lap
