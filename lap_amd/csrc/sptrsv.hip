// sptrsv.hip -- sparse triangular solves (hip_lower/upper_triangular_solve,
// initialize_and_analyze_L_and_U_solve).
//
// Level scheduling: the analysis assigns each row the length of its longest
// dependency chain (level) and sorts rows by level; a solve runs the levels
// in order, one launch per level, one lane per row (a workgroup per row for
// rows of more than kTriLongRow entries).  Each row's off-diagonal terms are
// subtracted in CSR order and the sum divided by the pivot, so every row is
// computed with exactly the reference's sequential arithmetic
// (simple_blas.c:46-85) and the result equals the sequential solve bitwise.
//
// Sync-free mode (option sptrsv_mode 1): one launch per solve.  Rows are cut
// into chunks of at most 64 rows of a single level, in level order.  A wave
// claims 4 consecutive chunks at a time from an atomic ticket and solves
// them in order, one lane per row.  Before the solve y is filled with a
// sentinel (a signalling-NaN bit pattern no arithmetic produces); a lane
// reads each dependency y[c] with 64-bit agent-scope atomic loads until it
// is not the sentinel, and stores its result with a 64-bit agent-scope
// atomic store, so the value is its own ready flag and no fence is needed.
// Progress: a lane only waits for rows of earlier levels, i.e. of earlier
// chunks, which waves that are already running have claimed (tickets are
// taken in order), and rows of one chunk never depend on each other.  Every
// wait is bounded: past the bound (far beyond any real wait) the lane raises
// an error flag, every other waiting lane gives up as soon as it sees it,
// and the host reports the failure, so the kernel always ends.  The
// arithmetic is the level-scheduled kernel's, in the same order: results
// are bitwise identical.
//
// Level-barrier mode (option sptrsv_mode 2): one persistent launch per
// solve, a grid-wide barrier between levels.  Every workgroup takes its
// share of the current level's rows (one lane per row), stores y with
// agent-scope stores (written through to the coherence point, so no cache
// write-back is needed), waits until its stores are complete, and arrives at
// a counter; wave 0 then polls the counter (wave-uniform, bounded) until all
// workgroups have arrived, and the next level's dependencies are read with
// agent-scope loads.  The grid is at most one workgroup per CU, so every
// workgroup is resident; every wait is bounded and raises the shared error
// flag when it runs out, so the kernel always ends.  Same row arithmetic as
// the level launches: bitwise identical.
#include <algorithm>

#include "lap_core.hpp"

namespace lap {

namespace {

// desc[t] = {row, first nonzero, end}: one load gives the row and its range
// (the row pointers are not read in the solve).
//
// One launch per level.  The level's short rows (the first `cnt` records)
// take one lane each; each of its `nlong` long rows (the records after them)
// takes a workgroup of its own: the workgroup forms the row's products a
// chunk at a time into LDS (two buffers: waves 1.. form the next chunk while
// lane 0 subtracts the current one in CSR order; the pivot's slot holds
// +0.0, and x - (+0.0) == x bit for bit), so a long row gets exactly the
// lane path's arithmetic while its loads proceed in parallel.  AT/VT
// (value/vector type) = double/double: the FP64 solve; float/float: FP32
// values and vectors with FP64 products and sums (the mixed-precision inner
// solve).
constexpr int kTriChunk = 2048;

// TREE (option sptrsv_long 1): a long row's products are summed by the
// workgroup as a fixed-order tree instead -- deterministic, no serial chain,
// but not the sequential order (tolerance parity, not bitwise).
template <class AT, class VT, bool TREE = false>
__global__ __launch_bounds__(kBlock) void k_tri_level(
    int cnt, int nlong, const int4 *__restrict__ desc, const int *__restrict__ ja,
    const AT *__restrict__ a, const double *__restrict__ diag, const VT *x, VT *y,
    const int *__restrict__ done) {
  if (done && *done) return;
  const int gshort = (cnt + kBlock - 1) / kBlock;
  if ((int)blockIdx.x < gshort) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= cnt) return;
    const int4 d = desc[t];
    const int row = d.x;
    double acc = (double)x[row];
    double piv = 0.0;
    bool in_row = false;
    const int e1 = d.z;
    for (int e = d.y; e < e1; ++e) {
      const int c = ja[e];
      if (c == row) {
        piv = (double)a[e];
        in_row = true;
        continue;
      }
      acc -= (double)a[e] * (double)y[c];
    }
    // diag is only passed (and only read) for matrices without in-row pivots
    if (!in_row && diag) piv = diag[row];
    y[row] = (VT)(acc / piv);
    return;
  }
  const int lr = (int)blockIdx.x - gshort;
  if (lr >= nlong) return;
  if (TREE) {
    __shared__ double red[kBlock / kWave];
    __shared__ int tpos;
    const int4 d = desc[cnt + lr];
    const int row = d.x;
    if (threadIdx.x == 0) tpos = -1;
    __syncthreads();
    double s = 0.0;
#pragma unroll 4
    for (int e = d.y + (int)threadIdx.x; e < d.z; e += kBlock) {
      const int c = ja[e];
      if (c == row)
        atomicMax(&tpos, e);
      else
        s += (double)a[e] * (double)y[c];
    }
    const double t = block_sum<kBlock>(s, red);  // ends with a barrier: tpos is final
    if (threadIdx.x == 0) {
      double piv = 0.0;
      if (tpos >= 0)
        piv = (double)a[tpos];
      else if (diag)
        piv = diag[row];
      y[row] = (VT)(((double)x[row] - t) / piv);
    }
    return;
  }
  __shared__ double prod[2][kTriChunk];
  __shared__ int dpos;
  const int4 d = desc[cnt + lr];
  const int row = d.x;
  const int nch = (d.z - d.y + kTriChunk - 1) / kTriChunk;
  if (threadIdx.x == 0) dpos = -1;
  __syncthreads();
  double acc = (double)x[row];
  // threads t0 .. t0 + nt - 1 form chunk ch's products, loads issued in
  // batches of four ahead of their use
  auto fill = [&](int ch, int t0, int nt) {
    const int base = d.y + ch * kTriChunk;
    const int m = min(kTriChunk, d.z - base);
    double *buf = prod[ch & 1];
    for (int k0 = (int)threadIdx.x - t0; k0 < m; k0 += 4 * nt) {
      int c[4];
      AT av[4];
      VT yv[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = k0 + j * nt;
        c[j] = k < m ? ja[base + k] : row;
        av[j] = k < m ? a[base + k] : AT(0);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) yv[j] = c[j] != row ? y[c[j]] : VT(0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = k0 + j * nt;
        if (k >= m) continue;
        if (c[j] == row) {
          buf[k] = 0.0;
          atomicMax(&dpos, base + k);  // the last pivot entry, as the lane path keeps
        } else {
          buf[k] = (double)av[j] * (double)yv[j];
        }
      }
    }
  };
  fill(0, 0, kBlock);
  // chunk ch is subtracted by lane 0 while waves 1.. form chunk ch + 1
  for (int ch = 0; ch < nch; ++ch) {
    __syncthreads();  // chunk ch formed, chunk ch - 1 consumed
    if (threadIdx.x < kWave) {
      if (threadIdx.x == 0)
        acc = lds_seq_sum<true>(acc, prod[ch & 1], 0, min(kTriChunk, d.z - d.y - ch * kTriChunk));
    } else if (ch + 1 < nch) {
      fill(ch + 1, kWave, kBlock - kWave);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double piv = 0.0;
    if (dpos >= 0)
      piv = (double)a[dpos];
    else if (diag)
      piv = diag[row];
    y[row] = (VT)(acc / piv);
  }
}

// Packed rows: lane t loads its row's w columns and values together (one
// dependent stage less than through desc), then the y of earlier levels.
// Same subtraction order and pivot as k_tri_level: bitwise identical.
template <int W>
__global__ __launch_bounds__(kBlock) void k_ell_pack(int n, const int4 *__restrict__ desc,
                                                     const int *__restrict__ ja,
                                                     const double *__restrict__ a,
                                                     int *__restrict__ eja,
                                                     double *__restrict__ ea) {
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
    const int4 d = desc[t];
    const int cnt = d.z - d.y;
#pragma unroll
    for (int k = 0; k < W; ++k) {
      eja[(size_t)t * W + k] = k < cnt ? ja[d.y + k] : -1;
      ea[(size_t)t * W + k] = k < cnt ? a[d.y + k] : 0.0;
    }
  }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_tri_level_ell(
    int cnt, const int *__restrict__ rows, const int *__restrict__ eja,
    const double *__restrict__ ea, const double *__restrict__ diag, const double *x, double *y,
    const int *__restrict__ done) {
  if (done && *done) return;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= cnt) return;
  const int row = rows[t];
  int c[W];
  double v[W];
#pragma unroll
  for (int k = 0; k < W; k += 4) {
    const int4 q = *reinterpret_cast<const int4 *>(eja + (size_t)t * W + k);
    c[k] = q.x;
    c[k + 1] = q.y;
    c[k + 2] = q.z;
    c[k + 3] = q.w;
  }
#pragma unroll
  for (int k = 0; k < W; k += 2) {
    const double2 q = *reinterpret_cast<const double2 *>(ea + (size_t)t * W + k);
    v[k] = q.x;
    v[k + 1] = q.y;
  }
  double acc = x[row];
  double piv = 0.0;
  bool in_row = false;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    if (c[k] < 0) break;
    if (c[k] == row) {
      piv = v[k];
      in_row = true;
      continue;
    }
    acc -= v[k] * y[c[k]];
  }
  if (!in_row && diag) piv = diag[row];
  y[row] = acc / piv;
}

constexpr unsigned long long kSentinel = 0x7ff44c4150534e54ull;  // sNaN payload
constexpr int kSfChunk = 64, kSfClaim = 4;
constexpr unsigned kSfSpinLimit = 1u << 20;

__device__ __forceinline__ double sf_load(const double *p) {
  const unsigned long long v = __hip_atomic_load(
      reinterpret_cast<const unsigned long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double((long long)v);
}

// y[i] = sentinel for i < n; state[0] (ticket) = 0
__global__ __launch_bounds__(kBlock) void k_sf_fill(int n, double *__restrict__ y, int *state,
                                                    const int *__restrict__ done) {
  if (done && *done) return;
  const size_t stride = (size_t)gridDim.x * kBlock;
  unsigned long long *yy = reinterpret_cast<unsigned long long *>(y);
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < (size_t)n; i += stride)
    yy[i] = kSentinel;
  if (blockIdx.x == 0 && threadIdx.x == 0) state[0] = 0;
}

// state[0]: chunk ticket; state[1]: error flag (dependency wait ran out)
__global__ __launch_bounds__(kBlock) void k_tri_syncfree(
    int nchunks, const int *__restrict__ chunk_ptr, const int *__restrict__ rows,
    const int *__restrict__ ia, const int *__restrict__ ja, const double *__restrict__ a,
    const double *__restrict__ diag, const double *__restrict__ x, double *y, int *state,
    const int *__restrict__ done, int naps) {
  if (done && *done) return;
  const int lane = threadIdx.x & 63;
  for (;;) {
    int first = 0;
    if (lane == 0) first = atomicAdd(&state[0], kSfClaim);
    first = __builtin_amdgcn_readfirstlane(__shfl(first, 0));
    if (first >= nchunks) break;
    const int last = min(first + kSfClaim, nchunks);
    for (int c = first; c < last; ++c) {
      const int t = chunk_ptr[c] + lane;
      if (t >= chunk_ptr[c + 1]) continue;
      const int row = rows[t];
      double acc = x[row];
      double piv = 0.0;
      bool in_row = false;
      const int e1 = ia[row + 1];
      for (int e = ia[row]; e < e1; ++e) {
        const int col = ja[e];
        if (col == row) {
          piv = a[e];
          in_row = true;
          continue;
        }
        double v = sf_load(y + col);
        unsigned spins = 0;
        while (__double_as_longlong(v) == (long long)kSentinel) {
          if (++spins > kSfSpinLimit ||
              __hip_atomic_load(&state[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            __hip_atomic_store(&state[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            v = __builtin_nan("");
            break;
          }
          for (int z = 0; z < naps; ++z) __builtin_amdgcn_s_sleep(8);  // ~512 cycles each
          v = sf_load(y + col);
        }
        acc -= a[e] * v;
      }
      if (!in_row && diag) piv = diag[row];
      const double r = acc / piv;
      __hip_atomic_store(reinterpret_cast<unsigned long long *>(y + row),
                         (unsigned long long)__double_as_longlong(r), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// state[2]: barrier arrivals (reset before every solve)
__global__ void k_tb_reset(int *state, const int *__restrict__ done) {
  if (done && *done) return;
  if (threadIdx.x == 0)
    __hip_atomic_store(&state[2], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr unsigned kTbSpinLimit = 1u << 22;

template <int BS>
__global__ __launch_bounds__(BS) void k_tri_barrier(
    int nlevels, const int *__restrict__ level_ptr, const int *__restrict__ rows,
    const int *__restrict__ ia, const int *__restrict__ ja, const double *__restrict__ a,
    const double *__restrict__ diag, const double *__restrict__ x, double *y, int *state,
    const int *__restrict__ done) {
  if (done && *done) return;
  __shared__ int failed;
  const int G = gridDim.x;
  for (int l = 0; l < nlevels; ++l) {
    const int l0 = level_ptr[l], l1 = level_ptr[l + 1];
    for (int t = l0 + blockIdx.x * BS + (int)threadIdx.x; t < l1; t += G * BS) {
      const int row = rows[t];
      double acc = x[row];
      double piv = 0.0;
      bool in_row = false;
      const int e1 = ia[row + 1];
      for (int e = ia[row]; e < e1; ++e) {
        const int c = ja[e];
        if (c == row) {
          piv = a[e];
          in_row = true;
          continue;
        }
        acc -= a[e] * sf_load(y + c);
      }
      if (!in_row && diag) piv = diag[row];
      const double r = acc / piv;
      __hip_atomic_store(reinterpret_cast<unsigned long long *>(y + row),
                         (unsigned long long)__double_as_longlong(r), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    if (l + 1 == nlevels) break;
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are complete
    __syncthreads();
    if (threadIdx.x < kWave) {  // wave 0: arrive, then wait for every workgroup
      if (threadIdx.x == 0) atomicAdd(&state[2], 1);
      const int target = (l + 1) * G;
      unsigned spins = 0;
      int ok = 1;
      for (;;) {
        const int cur = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(&state[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        const int err = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(&state[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (cur >= target) break;
        if (err || ++spins > kTbSpinLimit) {
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      if (threadIdx.x == 0) {
        if (!ok)
          __hip_atomic_store(&state[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        failed = !ok;
      }
    }
    __syncthreads();
    if (failed) break;
  }
}

int *sf_state() {  // [ticket, error, barrier arrivals, pad]
  static DevBufPtr buf;
  if (!buf) {
    buf = dev_alloc(4 * sizeof(int));
    if (buf) {
      LAP_HIP(hipMemsetAsync(buf->p, 0, 4 * sizeof(int), stream()));
      sync();
    }
  }
  return buf ? buf->as<int>() : nullptr;
}

using Key = std::tuple<const int *, const int *, int, int, bool>;
std::map<Key, std::unique_ptr<TriPlan>> g_plans;

}  // namespace

void release_tri_plans() { g_plans.clear(); }
void forget_tri_plans(const void *p) {
  for (auto it = g_plans.begin(); it != g_plans.end();) {
    if (std::get<0>(it->first) == p || std::get<1>(it->first) == p)
      it = g_plans.erase(it);
    else
      ++it;
  }
}

// Host-side level analysis of a triangular CSR pattern.
std::unique_ptr<TriPlan> tri_analyse_host(int n, const std::vector<int> &ia,
                                        const std::vector<int> &ja, bool lower) {
  auto p = std::make_unique<TriPlan>();
  p->n = n;
  p->nnz = n ? ia[n] : 0;
  p->lower = lower;
  std::vector<int> level((size_t)n, 0);
  int nlev = 0;
  bool has_diag = true;
  for (int s = 0; s < n; ++s) {
    const int i = lower ? s : n - 1 - s;
    int lv = 0;
    bool diag = false;
    for (int e = ia[i]; e < ia[i + 1]; ++e) {
      const int c = ja[e];
      if (c < 0 || c >= n) {
        set_error(LAP_ERR_STRUCT, "triangular solve: column %d out of range in row %d", c, i);
        return nullptr;
      }
      if (c == i) {
        diag = true;
        continue;
      }
      if ((lower && c > i) || (!lower && c < i)) {
        set_error(LAP_ERR_STRUCT, "%s triangular solve: entry (%d,%d) on the wrong side",
                  lower ? "lower" : "upper", i, c);
        return nullptr;
      }
      lv = std::max(lv, level[c] + 1);
    }
    has_diag = has_diag && diag;
    level[i] = lv;
    nlev = std::max(nlev, lv + 1);
  }
  if (n == 0) nlev = 0;
  p->nlevels = nlev;
  p->has_diag = has_diag;
  p->level_ptr.assign((size_t)nlev + 1, 0);
  for (int i = 0; i < n; ++i) p->level_ptr[level[i] + 1]++;
  for (int l = 0; l < nlev; ++l) p->level_ptr[l + 1] += p->level_ptr[l];
  // within a level: the rows of at most kTriLongRow entries, then the rest,
  // each in row order
  std::vector<int> nshort((size_t)nlev, 0);
  for (int i = 0; i < n; ++i)
    if (ia[i + 1] - ia[i] <= kTriLongRow) nshort[level[i]]++;
  p->level_long.resize((size_t)nlev);
  std::vector<int> fill(p->level_ptr.begin(), p->level_ptr.end() - 1), fill_long((size_t)nlev),
      order((size_t)n);
  for (int l = 0; l < nlev; ++l) fill_long[l] = p->level_long[l] = p->level_ptr[l] + nshort[l];
  for (int i = 0; i < n; ++i) {
    if (ia[i + 1] - ia[i] <= kTriLongRow)
      order[fill[level[i]]++] = i;
    else
      order[fill_long[level[i]]++] = i;
  }
  for (int l = 0; l < nlev; ++l)
    p->max_level_rows = std::max(p->max_level_rows, p->level_ptr[l + 1] - p->level_ptr[l]);
  for (int i = 0; i < n; ++i) p->max_row = std::max(p->max_row, ia[i + 1] - ia[i]);
  p->order = dev_upload(order);
  p->level_ptr_dev = dev_upload(p->level_ptr);
  std::vector<int> desc((size_t)n * 4);
  for (int t = 0; t < n; ++t) {
    desc[(size_t)t * 4] = order[(size_t)t];
    desc[(size_t)t * 4 + 1] = ia[order[(size_t)t]];
    desc[(size_t)t * 4 + 2] = ia[order[(size_t)t] + 1];
    desc[(size_t)t * 4 + 3] = 0;
  }
  p->desc = dev_upload(desc);
  // sync-free chunks: <= kSfChunk rows, never spanning two levels
  std::vector<int> cp;
  cp.reserve((size_t)n / kSfChunk + (size_t)nlev + 1);
  for (int l = 0; l < nlev; ++l)
    for (int t = p->level_ptr[l]; t < p->level_ptr[l + 1]; t += kSfChunk) cp.push_back(t);
  cp.push_back(n);
  p->nchunks = (int)cp.size() - 1;
  p->chunk_ptr = dev_upload(cp);
  return p;
}

const TriPlan *tri_plan(int n, int nnz, const int *ia, const int *ja, bool lower) {
  Key key{ia, ja, n, nnz, lower};
  auto it = g_plans.find(key);
  if (it != g_plans.end()) return it->second.get();
  if (n < 0 || nnz < 0) {
    set_error(LAP_ERR_ARG, "triangular solve: negative size");
    return nullptr;
  }
  std::vector<int> h_ia = dev_download(ia, (size_t)n + 1);
  const int total = n ? h_ia[n] : 0;
  if (n && (h_ia[0] != 0 || total < 0 || total > nnz)) {
    set_error(LAP_ERR_STRUCT, "triangular solve: row pointers inconsistent with nnz=%d", nnz);
    return nullptr;
  }
  std::vector<int> h_ja = dev_download(ja, (size_t)total);
  auto plan = tri_analyse_host(n, h_ia, h_ja, lower);
  if (!plan) return nullptr;
  TriPlan *raw = plan.get();
  g_plans[key] = std::move(plan);
  return raw;
}

int tri_error() {
  int *st = sf_state();
  if (!st) return 0;
  int h = 0;
  LAP_HIP(hipMemcpy(&h, st + 1, sizeof(int), hipMemcpyDeviceToHost));
  return h;
}

void tri_clear_error() {
  int *st = sf_state();
  if (st) LAP_HIP(hipMemsetAsync(st, 0, 4 * sizeof(int), stream()));
}

void tri_ell_build(const TriPlan &p, const int *ja, const double *a, TriEll &e) {
  e.w = 0;
  if (p.n <= 0 || p.max_row > 8) return;
  const int w = p.max_row <= 4 ? 4 : 8;
  e.ja = dev_alloc(sizeof(int) * (size_t)p.n * w);
  e.a = dev_alloc(sizeof(double) * (size_t)p.n * w);
  if (!e.ja || !e.a || !e.ja->p || !e.a->p) return;
  const unsigned g = grid_for((size_t)p.n, 1, kBlock, 65536);
  if (w == 4)
    k_ell_pack<4><<<g, kBlock, 0, stream()>>>(p.n, p.desc->as<int4>(), ja, a, e.ja->as<int>(),
                                              e.a->as<double>());
  else
    k_ell_pack<8><<<g, kBlock, 0, stream()>>>(p.n, p.desc->as<int4>(), ja, a, e.ja->as<int>(),
                                              e.a->as<double>());
  e.w = w;
}

void tri_solve_launch(const TriPlan &p, const int *ia, const int *ja, const double *a,
                      const double *diag, const double *x, double *y, const int *done,
                      const TriEll *ell) {
  const int *order = p.order->as<int>();
  if (ell && ell->w && options().sptrsv_mode == 0) {
    for (int l = 0; l < p.nlevels; ++l) {
      const int l0 = p.level_ptr[l], cnt = p.level_ptr[l + 1] - l0;
      const unsigned g = (unsigned)((cnt + kBlock - 1) / kBlock);
      const size_t off = (size_t)l0 * ell->w;
      if (ell->w == 4)
        k_tri_level_ell<4><<<g, kBlock, 0, stream()>>>(cnt, order + l0, ell->ja->as<int>() + off,
                                                       ell->a->as<double>() + off, diag, x, y,
                                                       done);
      else
        k_tri_level_ell<8><<<g, kBlock, 0, stream()>>>(cnt, order + l0, ell->ja->as<int>() + off,
                                                       ell->a->as<double>() + off, diag, x, y,
                                                       done);
    }
    return;
  }
  // in place (x == y) needs x intact while y is filled: level launches then
  if (options().sptrsv_mode == 2 && x != y && p.nlevels > 0) {
    int *st = sf_state();
    if (st) {
      k_tb_reset<<<1, kWave, 0, stream()>>>(st, done);
      const unsigned g = (unsigned)std::max(
          1, std::min((p.max_level_rows + kBlock - 1) / kBlock, device_cus()));
      k_tri_barrier<kBlock><<<g, kBlock, 0, stream()>>>(p.nlevels, p.level_ptr_dev->as<int>(),
                                                        order, ia, ja, a, diag, x, y, st, done);
      return;
    }
  }
  int *st = (options().sptrsv_mode == 1 && x != y) ? sf_state() : nullptr;
  if (st && p.n > 0) {
    const unsigned gf = grid_for((size_t)p.n, 4, kBlock, 2048);
    k_sf_fill<<<gf, kBlock, 0, stream()>>>(p.n, y, st, done);
    const int per_cu = std::max(1, std::min(16, options().sf_wg_per_cu));
    const unsigned g =
        (unsigned)std::max(1, std::min((p.nchunks + 3) / 4, per_cu * device_cus()));
    k_tri_syncfree<<<g, kBlock, 0, stream()>>>(p.nchunks, p.chunk_ptr->as<int>(), order, ia, ja,
                                               a, diag, x, y, st, done,
                                               std::max(1, options().sf_naps));
    return;
  }
  const int4 *desc = p.desc->as<int4>();
  for (int l = 0; l < p.nlevels; ++l) {
    const int cnt = p.level_long[l] - p.level_ptr[l], nlong = p.level_ptr[l + 1] - p.level_long[l];
    const unsigned g = (unsigned)((cnt + kBlock - 1) / kBlock + nlong);
    if (options().sptrsv_long == 1)
      k_tri_level<double, double, true><<<g, kBlock, 0, stream()>>>(
          cnt, nlong, desc + p.level_ptr[l], ja, a, diag, x, y, done);
    else
      k_tri_level<double, double><<<g, kBlock, 0, stream()>>>(cnt, nlong, desc + p.level_ptr[l],
                                                              ja, a, diag, x, y, done);
  }
}

template <class VT>
static void tri_mixed_launch(const TriPlan &p, const int *ja, const float *a, const VT *x, VT *y,
                             const int *done) {
  const int4 *desc = p.desc->as<int4>();
  for (int l = 0; l < p.nlevels; ++l) {
    const int cnt = p.level_long[l] - p.level_ptr[l], nlong = p.level_ptr[l + 1] - p.level_long[l];
    const unsigned g = (unsigned)((cnt + kBlock - 1) / kBlock + nlong);
    if (options().sptrsv_long == 1)
      k_tri_level<float, VT, true><<<g, kBlock, 0, stream()>>>(cnt, nlong, desc + p.level_ptr[l],
                                                               ja, a, nullptr, x, y, done);
    else
      k_tri_level<float, VT><<<g, kBlock, 0, stream()>>>(cnt, nlong, desc + p.level_ptr[l], ja,
                                                         a, nullptr, x, y, done);
  }
}

void tri_solve32_launch(const TriPlan &p, const int *ja, const float *a, const float *x, float *y,
                        const int *done) {
  tri_mixed_launch<float>(p, ja, a, x, y, done);
}


static void tri_entry(int n, int nnz, const int *ia, const int *ja, const double *a,
                      const double *diag, const double *x, double *y, bool lower) {
  if (n <= 0) return;
  const char *fn = lower ? "hip_lower_triangular_solve" : "hip_upper_triangular_solve";
  if (!check_device(fn, {{ia, "ia", sizeof(int) * ((size_t)n + 1)},
                         {ja, "ja", sizeof(int) * (size_t)nnz},
                         {a, "values", sizeof(double) * (size_t)nnz},
                         {x, "x", sizeof(double) * (size_t)n},
                         {y, "result", sizeof(double) * (size_t)n}}))
    return;
  const TriPlan *p = tri_plan(n, nnz, ia, ja, lower);
  if (!p) return;
  // With the pivots inside the matrix (the GPU drivers' L+D / D+U layout)
  // `diagonal` is ignored, as rocSPARSE ignores it (hip_blas.cpp:534-591);
  // the reference drivers may pass a stale host pointer there
  // (cg_driver2.c:246-248).
  const double *dg = nullptr;
  if (!p->has_diag) {
    if (!check_device(fn, {{diag, "diagonal (rows without an in-row pivot)",
                            sizeof(double) * (size_t)n}}))
      return;
    dg = diag;
  }
  tri_solve_launch(*p, ia, ja, a, dg, x, y, nullptr);
  sync();
  if (options().sptrsv_mode != 0 && tri_error()) {
    set_error(LAP_ERR_NUMERIC, "triangular solve: a dependency wait ran out (result invalid)");
    tri_clear_error();
  }
}

}  // namespace lap

using namespace lap;

extern "C" {

void initialize_and_analyze_L_and_U_solve(const int n, const int nnzL, int *lia, int *lja,
                                          double *la, const int nnzU, int *uia, int *uja,
                                          double *ua) {
  (void)la;
  (void)ua;
  if (n <= 0) return;
  if (!check_device("initialize_and_analyze_L_and_U_solve",
                    {{lia, "lia"}, {lja, "lja"}, {uia, "uia"}, {uja, "uja"}}))
    return;
  g_plans.erase(Key{lia, lja, n, nnzL, true});  // explicit analysis rebuilds
  g_plans.erase(Key{uia, uja, n, nnzU, false});
  tri_plan(n, nnzL, lia, lja, true);
  tri_plan(n, nnzU, uia, uja, false);
}

void hip_lower_triangular_solve(const int n, const int nnzL, const int *lia, const int *lja,
                                const double *la, const double *diagonal, const double *x,
                                double *result) {
  tri_entry(n, nnzL, lia, lja, la, diagonal, x, result, true);
}

void hip_upper_triangular_solve(const int n, const int nnzU, const int *uia, const int *uja,
                                const double *ua, const double *diagonal, const double *x,
                                double *result) {
  tri_entry(n, nnzU, uia, uja, ua, diagonal, x, result, false);
}

void lower_triangular_solve(const int n, const int nnz, const int *lia, const int *lja,
                            const double *la, const double *diag, const double *x,
                            double *result) {
  hip_lower_triangular_solve(n, nnz, lia, lja, la, diag, x, result);
}

void upper_triangular_solve(const int n, const int nnz, const int *uia, const int *uja,
                            const double *ua, const double *diag, const double *x,
                            double *result) {
  hip_upper_triangular_solve(n, nnz, uia, uja, ua, diag, x, result);
}

}  // extern "C"
