// spmv.hip -- CSR SpMV y = alpha*A*x + beta*y (hip_csr_matvec, analyze_spmv).
//
// Design (HBM-bound, 12 B per nonzero + 20 B per row):
//  * Analysis (host, once per (ia, ja)) cuts the rows into row blocks of at
//    most `spmv_rows` rows and kTile nonzeros.  A row longer than kTile gets
//    a block of its own and is summed by the whole workgroup (tree sum).
//  * Short-row blocks ("stream" mode): the 256 lanes read the block's
//    nonzeros coalesced (ja, a) and gather x, writing alpha*a*x products to
//    LDS; then one lane per row sums its products from LDS in CSR order.
//    So every row is summed sequentially in CSR order, exactly as the
//    reference's simple_csr_matvec (simple_blas.c:24-44), while all HBM
//    traffic for the matrix is fully coalesced.
//  * Blocks are remapped so that each XCD owns a contiguous range of rows
//    (x entries shared by neighbouring rows stay in one XCD's L2).
#include <algorithm>
#include <cstring>

#include "lap_core.hpp"
#include "spmv_core.hpp"

namespace lap {

namespace {

struct PlainEpi {
  double *y;
  double beta;
  __device__ double init(int row) const { return beta == 0.0 ? 0.0 : y[row] * beta; }
  __device__ void store(int row, double v) const { y[row] = v; }
};

template <int BS, int TILE>
__global__ __launch_bounds__(BS) void k_spmv(int nblk, const int *__restrict__ rowblk,
                                             const int *__restrict__ ia,
                                             const int *__restrict__ ja,
                                             const double *__restrict__ a,
                                             const double *__restrict__ x, double *y,
                                             double alpha, double beta, int xcd,
                                             const int *done) {
  if (done && *done) return;
  __shared__ double prod[TILE];
  __shared__ double red[BS / kWave];
  PlainEpi epi{y, beta};
  spmv_walk<BS, TILE>(block_walk(nblk, xcd), rowblk, ia, ja, a, x, alpha, prod, red, epi);
}

using Key = std::tuple<const int *, const int *, int, int, int>;
std::map<Key, std::unique_ptr<SpmvPlan>> g_plans;

}  // namespace

const SpmvPlan *spmv_plan(int n, int nnz, const int *ia, const int *ja) {
  const int R = options().spmv_rows;
  Key key{ia, ja, n, nnz, R};
  auto it = g_plans.find(key);
  if (it != g_plans.end()) return it->second.get();
  if (n < 0 || nnz < 0) {
    set_error(LAP_ERR_ARG, "spmv: negative size n=%d nnz=%d", n, nnz);
    return nullptr;
  }
  std::vector<int> h_ia = dev_download(ia, (size_t)n + 1);
  if (n > 0 && (h_ia[0] != 0 || h_ia[n] > nnz)) {
    set_error(LAP_ERR_STRUCT, "spmv: row pointers inconsistent (ia[0]=%d ia[n]=%d nnz=%d)",
              h_ia[0], h_ia[n], nnz);
    return nullptr;
  }
  auto plan = std::make_unique<SpmvPlan>();
  plan->n = n;
  plan->nnz = nnz;
  std::vector<int> starts;
  starts.reserve((size_t)n / (R > 0 ? R : 1) + 2);
  starts.push_back(0);
  int r = 0, max_row = 0;
  while (r < n) {
    const int len = h_ia[r + 1] - h_ia[r];
    if (len < 0) {
      set_error(LAP_ERR_STRUCT, "spmv: row pointers decrease at row %d", r);
      return nullptr;
    }
    max_row = std::max(max_row, len);
    if (len > kSpmvTile) {  // long row: a block of its own
      starts.push_back(++r);
      continue;
    }
    int r1 = r + 1;
    while (r1 < n && r1 - r < R && h_ia[r1 + 1] - h_ia[r] <= kSpmvTile) {
      max_row = std::max(max_row, h_ia[r1 + 1] - h_ia[r1]);
      ++r1;
    }
    starts.push_back(r1);
    r = r1;
  }
  // every column index must address x: checked once here, so no launch on
  // this plan can read outside x
  if (n > 0 && h_ia[n] > 0) {
    const std::vector<int> h_ja = dev_download(ja, (size_t)h_ia[n]);
    for (size_t e = 0; e < h_ja.size(); ++e)
      if (h_ja[e] < 0 || h_ja[e] >= n) {
        set_error(LAP_ERR_STRUCT, "spmv: column index %d out of range at nonzero %zu",
                  h_ja[e], e);
        return nullptr;
      }
  }
  plan->nblocks = (int)starts.size() - 1;
  plan->max_row = max_row;
  plan->rowblk = dev_upload(starts);
  SpmvPlan *raw = plan.get();
  g_plans[key] = std::move(plan);
  return raw;
}

unsigned spmv_grid(const SpmvPlan &p) {
  const int cap = options().spmv_wg_per_cu * device_cus();
  return (unsigned)std::max(1, std::min(p.nblocks, cap));
}

void spmv_launch(const SpmvPlan &p, const int *ia, const int *ja, const double *a,
                 const double *x, double *y, double alpha, double beta, const int *done) {
  if (p.nblocks == 0) return;
  const unsigned g = spmv_grid(p);
  const int xcd = options().spmv_walk;
  const int *rb = p.rowblk->as<int>();
  k_spmv<kBlock, kSpmvTile><<<g, kBlock, 0, stream()>>>(p.nblocks, rb, ia, ja, a, x, y, alpha,
                                                     beta, xcd, done);
}

void release_tri_plans();
void release_ic_plans();
void release_mix_plans();
void release_batch_plans();
void release_block_plans();
void forget_tri_plans(const void *p);
void forget_ic_plans(const void *p);
void forget_mix_plans(const void *p);
void forget_batch_plans(const void *p);
void forget_block_plans(const void *p);
void forget_pair_plans(const void *p);
void release_pair_plans();

void forget_plans(const void *p) {
  for (auto it = g_plans.begin(); it != g_plans.end();) {
    if (std::get<0>(it->first) == p || std::get<1>(it->first) == p)
      it = g_plans.erase(it);
    else
      ++it;
  }
  forget_tri_plans(p);
  forget_ic_plans(p);
  forget_mix_plans(p);
  forget_batch_plans(p);
  forget_block_plans(p);
  forget_pair_plans(p);
}
void release_plans() {
  release_pair_plans();
  release_block_plans();
  release_batch_plans();
  release_mix_plans();
  g_plans.clear();
  release_tri_plans();
  release_ic_plans();
}

}  // namespace lap

using namespace lap;

extern "C" {

void analyze_spmv(const int n, const int nnz, int *ia, int *ja, double *a, const double *x,
                  double *result, char *option) {
  (void)a;
  (void)x;
  (void)result;
  if (option && strcmp(option, "A") && strcmp(option, "L") && strcmp(option, "U"))
    fprintf(stderr, "[lapgfx] analyze_spmv: unknown option '%s' (analysed anyway)\n", option);
  if (n <= 0 || !check_device("analyze_spmv", {{ia, "ia", sizeof(int) * ((size_t)n + 1)},
                                               {ja, "ja", sizeof(int) * (size_t)nnz}}))
    return;
  // an explicit analysis always rebuilds (the arrays may hold a new matrix)
  g_plans.erase(Key{ia, ja, n, nnz, options().spmv_rows});
  spmv_plan(n, nnz, ia, ja);
}

void hip_csr_matvec(const int n, const int nnz, const int *ia, const int *ja, const double *a,
                    const double *x, double *result, const double *al, const double *bet,
                    const char *kind) {
  (void)kind;  // plans are keyed by the arrays themselves
  if (n <= 0) return;
  if (!check_device("hip_csr_matvec",
                    {{ia, "ia", sizeof(int) * ((size_t)n + 1)}, {ja, "ja", sizeof(int) * (size_t)nnz},
                     {a, "values", sizeof(double) * (size_t)nnz}, {x, "x", sizeof(double) * (size_t)n},
                     {result, "result", sizeof(double) * (size_t)n}}))
    return;
  const SpmvPlan *p = spmv_plan(n, nnz, ia, ja);
  if (!p) return;
  spmv_launch(*p, ia, ja, a, x, result, *al, *bet);
  sync();
}

void csr_matvec(const int n, const int nnz, const int *ia, const int *ja, const double *a,
                const double *x, double *result, const double *al, const double *bet,
                const char *kind) {
  hip_csr_matvec(n, nnz, ia, ja, a, x, result, al, bet, kind);
}

void lap_forget(const void *p) {
  sync();
  forget_plans(p);
}

double lap_spmv_timed(int n, int nnz, const int *ia, const int *ja, const double *a,
                      const double *x, double *y, int reps) {
  if (n <= 0 || reps < 1 ||
      !check_device("lap_spmv_timed", {{ia, "ia"}, {ja, "ja"}, {a, "a"}, {x, "x"}, {y, "y"}}))
    return -1.0;
  const SpmvPlan *p = spmv_plan(n, nnz, ia, ja);
  if (!p) return -1.0;
  spmv_launch(*p, ia, ja, a, x, y, 1.0, 0.0);  // warm
  hipEvent_t e0, e1;
  LAP_HIP(hipEventCreate(&e0));
  LAP_HIP(hipEventCreate(&e1));
  LAP_HIP(hipEventRecord(e0, stream()));
  for (int i = 0; i < reps; ++i) spmv_launch(*p, ia, ja, a, x, y, 1.0, 0.0);
  LAP_HIP(hipEventRecord(e1, stream()));
  LAP_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  LAP_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / reps;
}

}  // extern "C"
