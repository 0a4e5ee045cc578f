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
// Chunked mode (option sptrsv_mode 3, the default): one persistent launch
// per solve, one workgroup per chunk of rows (all workgroups resident: at
// most one per CU).  A chunk runs its rows level by level; a level of a chunk
// is a "step", one lane per row, the rows packed as fixed-size records.
// Every dependency of a step lies in an earlier step of its own chunk (LDS
// ring or y) or in another chunk, at a lower level; before a step, compute
// wave 0 makes sure every chunk the step depends on has published (in its
// progress counter) the highest level the step needs from it --
// point-to-point waits, listed per step by the analysis and pruned to the
// ones not already implied by an earlier wait of the chunk.  y is stored with
// agent-scope stores and dependencies are read with agent-scope loads (both
// coherent across the XCDs' L2s without cache write-backs); a step's stores
// are complete (s_waitcnt) before its level is published.  Every wait is
// bounded and raises the shared error flag when it runs out, so the kernel
// always ends.  Same row arithmetic as the level launches: bitwise
// identical.  Triangles whose rows do not fit a record keep the level
// launches (option sptrsv_mode 0 forces them everywhere).
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <thread>

#include "lap_core.hpp"

namespace lap {

namespace {

// desc[t] = {row, first nonzero, end, pivot}: one load gives the row, its
// range and the position of its last diagonal entry (-1: none) (the row
// pointers are not read in the solve).
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
constexpr int kTriFill = 12;  // products per thread per fill batch (2048 / 192 <= 12)

__device__ __forceinline__ double sf_load(const double *p) {
  const unsigned long long v = __hip_atomic_load(
      reinterpret_cast<const unsigned long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double((long long)v);
}
__device__ __forceinline__ float sf_load(const float *p) {
  const unsigned v = __hip_atomic_load(reinterpret_cast<const unsigned *>(p), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return __uint_as_float(v);
}
__device__ __forceinline__ void sf_store(double *p, double v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long *>(p),
                     (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void sf_store(float *p, float v) {
  __hip_atomic_store(reinterpret_cast<unsigned *>(p), __float_as_uint(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// s - p[0] - p[1] - ... - p[m - 1], one term at a time in that order (the
// sequential loop's result bit for bit).  p is LDS, 16-byte aligned, with at
// least 12 readable elements past m (their values are never used).  Batches
// of 12 terms alternate between two register sets and the next batch's loads
// are issued before the current batch's subtractions, so the dependent chain
// pays the FP64 add latency (~6.5 cycles) per term rather than an LDS round
// trip per batch; at most 12 LDS loads are in flight (the counter holds 15).
__device__ __forceinline__ double lds_seq_sub_ahead(double s, const double *p, int m) {
  constexpr int H = 6;  // 16-byte loads per batch
  const double2 *q = reinterpret_cast<const double2 *>(p);
  const int full = m / (2 * H);
  double2 v[H], w[H];
#pragma unroll
  for (int j = 0; j < H; ++j) v[j] = q[j];
  int b = 0;
  // (the scheduling barriers keep the compiler from gathering both batches'
  // loads in one place, which would expose a round trip again)
  for (; b + 2 <= full; b += 2) {
#pragma unroll
    for (int j = 0; j < H; ++j) w[j] = q[(b + 1) * H + j];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < H; ++j) {
      s -= v[j].x;
      s -= v[j].y;
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < H; ++j) v[j] = q[(b + 2) * H + j];  // ends <= m + 11
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < H; ++j) {
      s -= w[j].x;
      s -= w[j].y;
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  if (b < full) {
#pragma unroll
    for (int j = 0; j < H; ++j) {
      s -= v[j].x;
      s -= v[j].y;
    }
    ++b;
  }
  for (int k = b * 2 * H; k < m; ++k) s -= p[k];
  return s;
}

// TREE (option sptrsv_long 1): a long row's products are summed in a fixed
// tree order instead -- deterministic, no serial chain, but not the
// sequential order (tolerance parity, not bitwise).  The row is split into
// segments of kTriSeg entries, one workgroup each (the launch's long-row
// workgroups index segd: {level position, segment, segments of the row, the
// row's slot}); each forms its segment's sum as a workgroup tree, and a row
// of several segments is finished by the workgroup that arrives last on the
// row's ticket: it adds the segments' sums in segment order (so the result
// does not depend on which one that is) and resets the ticket for the next
// solve.  Partial sums and the ticket are agent-scope, ordered by fences.
template <class AT, class VT, bool TREE = false>
__global__ __launch_bounds__(kBlock) void k_tri_level(
    int cnt, int nlong, const int4 *__restrict__ desc, const int *__restrict__ ja,
    const AT *__restrict__ a, const double *__restrict__ diag, const VT *x, VT *y,
    const int *__restrict__ done, const int4 *__restrict__ segd, double *part, int *tick) {
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
    __shared__ int last;
    const int4 sd = segd[lr];
    const int4 d = desc[sd.x];
    const int row = d.x;
    const int e0 = d.y + sd.y * kTriSeg, e1 = min(d.z, e0 + kTriSeg);
    constexpr int B = kTriSeg / kBlock;
    int c[B];
    AT av[B];
    VT yv[B];
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int e = e0 + (int)threadIdx.x + j * kBlock;
      c[j] = e < e1 ? ja[e] : row;
      av[j] = e < e1 ? a[e] : AT(0);
    }
#pragma unroll
    for (int j = 0; j < B; ++j) yv[j] = c[j] != row ? y[c[j]] : VT(0);
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < B; ++j)
      if (c[j] != row) s += (double)av[j] * (double)yv[j];
    const double t = block_sum<kBlock>(s, red);  // valid in thread 0
    auto finish = [&](double sum) {
      const double piv = d.w >= 0 ? (double)a[d.w] : diag ? diag[row] : 0.0;
      y[row] = (VT)(((double)x[row] - sum) / piv);
    };
    if (sd.z == 1) {
      if (threadIdx.x == 0) finish(t);
      return;
    }
    if (threadIdx.x == 0) {
      sf_store(part + sd.w + sd.y, t);
      __threadfence();
      last = atomicAdd(tick + sd.w, 1) == sd.z - 1;
    }
    __syncthreads();
    if (threadIdx.x != 0 || !last) return;
    __threadfence();
    double sum = 0.0;
    for (int j = 0; j < sd.z; ++j) sum += sf_load(part + sd.w + j);
    __hip_atomic_store(tick + sd.w, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    finish(sum);
    return;
  }
  __shared__ __align__(16) double prod[2][kTriChunk + 16];  // + lds_seq_sub_ahead's read-ahead
  __shared__ int dpos;
  const int4 d = desc[cnt + lr];
  const int row = d.x;
  const int nch = (d.z - d.y + kTriChunk - 1) / kTriChunk;
  if (threadIdx.x == 0) dpos = -1;
  __syncthreads();
  double acc = (double)x[row];
  // threads t0 .. t0 + nt - 1 form chunk ch's products, all of a thread's
  // loads for the chunk issued together (one batch of kTriFill)
  auto fill = [&](int ch, int t0, int nt) {
    const int base = d.y + ch * kTriChunk;
    const int m = min(kTriChunk, d.z - base);
    double *buf = prod[ch & 1];
    for (int k0 = (int)threadIdx.x - t0; k0 < m; k0 += kTriFill * nt) {
      int c[kTriFill];
      AT av[kTriFill];
      VT yv[kTriFill];
#pragma unroll
      for (int j = 0; j < kTriFill; ++j) {
        const int k = k0 + j * nt;
        c[j] = k < m ? ja[base + k] : row;
        av[j] = k < m ? a[base + k] : AT(0);
      }
#pragma unroll
      for (int j = 0; j < kTriFill; ++j) yv[j] = c[j] != row ? y[c[j]] : VT(0);
#pragma unroll
      for (int j = 0; j < kTriFill; ++j) {
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
        acc = lds_seq_sub_ahead(acc, prod[ch & 1], min(kTriChunk, d.z - d.y - ch * kTriChunk));
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

// ---------------------------------------------------------- chunked mode
constexpr int kChMinRows = 2048;  // rows per chunk at least
constexpr int kChMax = 1024;      // chunks at most (the kernel's LDS table)
constexpr int kChStepRows = kBlock;  // rows per step of a chunk at most (one per lane)
constexpr int kChRing = 4;        // steps of results kept in LDS
constexpr int kChSteps = 512;     // steps of a chunk at most (LDS table)
constexpr int kChWaits = 1024;    // waits of a chunk at most (LDS table)
constexpr unsigned kChSpinLimit = 1u << 22;  // polls of one wait before the solve fails

__global__ __launch_bounds__(kBlock) void k_ch_reset(int nchunks, int *progress,
                                                     const int *__restrict__ done) {
  if (done && *done) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nchunks; i += gridDim.x * kBlock)
    __hip_atomic_store(&progress[i], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroup barrier for LDS traffic only: waits for this wave's LDS
// operations, not for its outstanding global loads (__syncthreads would wait
// for every memory operation of the wave).  The signal fences keep the
// compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_s_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// s_waitcnt encoding (gfx9) for vmcnt(N) alone: lgkmcnt and expcnt at their
// maxima.
constexpr unsigned vmcnt_only(unsigned n) {
  return (n & 15u) | ((n >> 4) << 14) | (7u << 4) | (15u << 8);
}

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void global_void_t;

// The chunked schedule's records, one per row in schedule order (bytes):
// [0, 4W) the off-diagonal dependency codes in CSR order (-1 padding, >= 0 a
// row of y, <= -2 the LDS ring slot -(code + 2)); [4W, 4W + 4) the row; then
// the off-diagonal values and the pivot: FP64 values at [4W + 16, 12W + 16)
// and [12W + 16, 12W + 24), FP32 values at [4W + 8, 8W + 8) and
// [8W + 8, 8W + 12).  The size is a multiple of 16 bytes.
template <int W, class AT>
struct ChRec;
template <int W>
struct ChRec<W, double> {
  static constexpr int bytes = 12 * W + 32;
  static constexpr int row = W;                  // int index
  static constexpr int val = (4 * W + 16) / 8;   // value index
  static constexpr int piv = (12 * W + 16) / 8;  // value index
  static constexpr int dma = bytes / 4;          // 16-byte loads of 64 lanes per kBlock records
};
template <int W>
struct ChRec<W, float> {
  static constexpr int bytes = (8 * W + 12 + 15) / 16 * 16;
  static constexpr int row = W;          // int index
  static constexpr int val = W + 2;      // value index
  static constexpr int piv = 2 * W + 2;  // value index
  static constexpr int dma = bytes / 4;
};

// One workgroup per chunk: kBlock compute lanes (one row each per step) and
// one prefetch wave.  progress[c]: highest level of chunk c whose rows are
// all complete and visible; state[1]: error flag (a wait ran out).
//  * the prefetch wave copies each step's records and right-hand sides (xc:
//    x in schedule order, gathered beforehand by k_ch_gather) into an LDS
//    slot D steps ahead with direct global-to-LDS loads; before each step
//    ends it waits, by count, for the next step's slot only (its loads are
//    all of one kind and complete in order).  No compute lane ever waits on
//    a copy;
//  * a step starts with compute wave 0 checking its waits against a table
//    of the levels already seen published, which that wave refreshes by
//    polling the next step's producers while the current step runs.  Only
//    when a wait is not covered does the chunk block: it first publishes
//    everything it has completed, so that no chunk waits on finished but
//    unpublished work, and polls until the levels it needs are published;
//  * each row reads rows of this chunk from the last kChRing steps out of
//    the LDS ring and the others from y (agent-scope loads), and writes its
//    result to the ring and to y (agent-scope store).  A step ends with the
//    lanes waiting, by count, for the previous step's store only and a
//    barrier, after which the previous step's level is published: one step
//    behind, so that no lane waits on a store it has just issued.
// AT: the values' type, VT: the vectors' (FP32 values and vectors are
// summed in FP64 and the result rounded once, as the FP32 level launches do).
template <class AT, class VT, int W, int D, int NS>
__global__ __launch_bounds__(kBlock + kWave) void k_tri_chunks(
    const int *__restrict__ cstep, const int *__restrict__ sptr, const int *__restrict__ slev,
    const int *__restrict__ wptr, const int *__restrict__ wchk, const int *__restrict__ wlev,
    const char *__restrict__ rec, const VT *xc, VT *y,
    int *progress, int *state, const int *__restrict__ done) {
  if (done && *done) return;
  using R = ChRec<W, AT>;
  constexpr int XB = kBlock * (int)sizeof(VT);  // the step's right-hand sides
  constexpr int SLOT = kBlock * R::bytes + XB;
  constexpr int PER = R::dma + XB / (kWave * 4);  // loads per step: records, right-hand sides
  static_assert((D - 1) * PER <= 63, "vmcnt range");
  static_assert(D < NS, "a slot in use is never refilled");
  __shared__ __attribute__((aligned(16))) char slots[NS * SLOT];
  __shared__ int seen[kChMax];  // highest level seen published, per chunk
  __shared__ VT ring[kChRing * kBlock];
  __shared__ int tsptr[kChSteps + 1], tslev[kChSteps], twptr[kChSteps + 1];
  __shared__ int twchk[kChWaits], twlev[kChWaits];
  __shared__ int failed, must_wait;
  const int c = blockIdx.x;
  const bool pf = threadIdx.x >= kBlock;  // the prefetch wave
  const int p = threadIdx.x;              // a compute lane's row position in a step
  const int k0 = __builtin_amdgcn_readfirstlane(cstep[c]);
  const int ns = __builtin_amdgcn_readfirstlane(cstep[c + 1]) - k0;
  const int q0 = wptr[k0];
  const int nw = wptr[k0 + ns] - q0;
  constexpr int nt = kBlock + kWave;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += nt) seen[i] = -1;
  for (int i = threadIdx.x; i <= ns; i += nt) {
    tsptr[i] = sptr[k0 + i];
    twptr[i] = wptr[k0 + i] - q0;
    if (i < ns) tslev[i] = slev[k0 + i];
  }
  for (int i = threadIdx.x; i < nw; i += nt) {
    twchk[i] = wchk[q0 + i];
    twlev[i] = wlev[q0 + i];
  }
  if (threadIdx.x == 0) {
    failed = 0;
    must_wait = 0;
  }
  __syncthreads();
  const int tfirst = tsptr[0];
  const int lane = (int)threadIdx.x - kBlock;  // prefetch wave lane
  // prefetch wave: step rel's records into its slot (for rel >= ns a
  // harmless copy of the chunk's first ones, so every step issues as many)
  auto copy_records = [&](int rel) {
    const int t0 = rel < ns ? tsptr[rel] : tfirst;
    char *slot = slots + (rel % NS) * SLOT;
    const char *src = rec + (size_t)t0 * R::bytes + lane * 16;
#pragma unroll
    for (int i = 0; i < R::dma; ++i)
      __builtin_amdgcn_global_load_lds((global_void_t *)(src + i * 1024),
                                       (lds_void_t *)(slot + i * 1024), 16, 0, 0);
    const char *xs = reinterpret_cast<const char *>(xc + t0) + lane * 4;
#pragma unroll
    for (int i = 0; i < XB / (kWave * 4); ++i)
      __builtin_amdgcn_global_load_lds((global_void_t *)(xs + i * 256),
                                       (lds_void_t *)(slot + kBlock * R::bytes + i * 256), 4, 0,
                                       0);
  };
  if (pf) {
#pragma unroll
    for (int r = 0; r < D; ++r) copy_records(r);
    __builtin_amdgcn_s_waitcnt(vmcnt_only((D - 1) * PER));  // step 0's slot
  }
  for (int rel = 0; rel < ns; ++rel) {
    if (threadIdx.x < kWave) {  // compute wave 0: are the step's waits already satisfied?
      bool ready = true;
      for (int q = twptr[rel] + (int)threadIdx.x; q < twptr[rel + 1]; q += kWave)
        ready = ready && seen[twchk[q]] >= twlev[q];
      const bool all = __builtin_amdgcn_ballot_w64(!ready) == 0;
      if (threadIdx.x == 0) must_wait = !all;
    }
    lds_barrier();  // must_wait set; the previous step's ring entries and slot done
    if (__builtin_amdgcn_readfirstlane(must_wait)) {
      // about to block: first publish the previous step (its stores complete),
      // so that no chunk ever waits on work that is done but unpublished
      if (!pf) __builtin_amdgcn_s_waitcnt(0);
      lds_barrier();
      if (threadIdx.x == 0 && rel > 0)
        __hip_atomic_store(&progress[c], tslev[rel - 1], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      if (threadIdx.x < kWave) {  // wave 0: the step's waits, one lane per producer
        int ok = 1;
        const int qe = twptr[rel + 1];
        for (int q = twptr[rel] + (int)threadIdx.x; q < qe; q += kWave) {
          const int pc = twchk[q], need = twlev[q];
          if (seen[pc] >= need) continue;
          unsigned spins = 0;
          int v = __hip_atomic_load(&progress[pc], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          while (v < need) {
            if (++spins > kChSpinLimit ||
                __hip_atomic_load(&state[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
              ok = 0;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            v = __hip_atomic_load(&progress[pc], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
          seen[pc] = v;
          if (!ok) break;
        }
        if (!ok) {
          __hip_atomic_store(&state[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          failed = 1;
        }
      }
      lds_barrier();
      if (__builtin_amdgcn_readfirstlane(failed)) break;
    }
    if (pf) {
      copy_records(rel + D);
      // the next step's slot is complete; later ones may still be in flight
      __builtin_amdgcn_s_waitcnt(vmcnt_only((D - 1) * PER));
    } else {
      // wave 0: poll a producer of the next step not yet seen far enough,
      // issued before this step's loads and read after them
      int poll_pc = -1, poll_v = 0;
      if (threadIdx.x < kWave && rel + 1 < ns) {
        const int q = twptr[rel + 1] + (int)threadIdx.x;
        if (q < twptr[rel + 2] && seen[twchk[q]] < twlev[q]) {
          poll_pc = twchk[q];
          poll_v = __hip_atomic_load(&progress[poll_pc], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      const bool has_row = tsptr[rel] + p < tsptr[rel + 1];
      int row = 0;
      VT r = VT(0);
      if (has_row) {
        const char *slot = slots + (rel % NS) * SLOT;
        const int *ri = reinterpret_cast<const int *>(slot + p * R::bytes);
        const AT *rd = reinterpret_cast<const AT *>(slot + p * R::bytes);
        int cc[W];
        VT yv[W];
#pragma unroll
        for (int j = 0; j < W; ++j) cc[j] = ri[j];
#pragma unroll
        for (int j = 0; j < W; ++j)  // the dependencies on y first: one wait for all
          yv[j] = cc[j] >= 0 ? sf_load(y + cc[j]) : VT(0);
        row = ri[R::row];
        const VT xr = reinterpret_cast<const VT *>(slot + kBlock * R::bytes)[p];
        const double piv = (double)rd[R::piv];
#pragma unroll
        for (int j = 0; j < W; ++j)
          if (cc[j] < -1) yv[j] = ring[-(cc[j] + 2)];
        double acc = (double)xr;
#pragma unroll
        for (int j = 0; j < W; ++j) {
          if (cc[j] == -1) break;
          acc -= (double)rd[R::val + j] * (double)yv[j];
        }
        r = (VT)(acc / piv);
        ring[(rel % kChRing) * kBlock + p] = r;
      }
      if (poll_pc >= 0) atomicMax(&seen[poll_pc], poll_v);
      if (has_row) {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        sf_store(y + row, r);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
      }
      // The loads of this step were consumed above, so only stores are in
      // flight, and stores complete in order: all but this step's store
      // (present when one of the wave's lanes has a row) are complete, the
      // previous step's included.
      if (__builtin_amdgcn_ballot_w64(has_row) != 0)
        __builtin_amdgcn_s_waitcnt(vmcnt_only(1));
      else
        __builtin_amdgcn_s_waitcnt(vmcnt_only(0));
    }
    lds_barrier();  // ring entries written; every wave's previous-step stores complete
    if (threadIdx.x == 0 && rel > 0)
      __hip_atomic_store(&progress[c], tslev[rel - 1], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
  // the last step: its stores complete, then its level published
  if (!pf) __builtin_amdgcn_s_waitcnt(0);
  lds_barrier();
  if (threadIdx.x == 0 && ns > 0 && !__builtin_amdgcn_readfirstlane(failed))
    __hip_atomic_store(&progress[c], tslev[ns - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (pf) __builtin_amdgcn_s_waitcnt(0);  // no copy into LDS outlives the workgroup
}

// xc[t] = x[row of schedule position t]: the right-hand side in the chunked
// schedule's order, for k_tri_chunks' prefetch.
// The rows of neighbouring schedule positions share lines of x, so each XCD
// (workgroups are dealt to the 8 XCDs round-robin) sweeps one contiguous
// eighth of the positions: its workgroups then read the same lines out of
// their own L2.  The grid is a multiple of 8.
template <class VT>
__global__ __launch_bounds__(kBlock) void k_ch_gather(int n, const int *__restrict__ crow,
                                                      const VT *__restrict__ x,
                                                      VT *__restrict__ xc,
                                                      const int *__restrict__ done) {
  if (done && *done) return;
  const int xcd = blockIdx.x & 7, per = gridDim.x >> 3;
  const long long lo = (long long)n * xcd / 8, hi = (long long)n * (xcd + 1) / 8;
  for (long long t = lo + (long long)(blockIdx.x >> 3) * kBlock + threadIdx.x; t < hi;
       t += (long long)per * kBlock)
    st_nt(xc + t, x[ld_nt(crow + t)]);  // the map and xc are streamed; x lines are reused
}

// The records of the chunked schedule for one set of values (k_tri_chunks).
template <int W, class AT>
__global__ __launch_bounds__(kBlock) void k_ch_pack(int n, const int *__restrict__ cmap,
                                                    const int4 *__restrict__ desc,
                                                    const int *__restrict__ ja,
                                                    const AT *__restrict__ a,
                                                    const int *__restrict__ code,
                                                    char *__restrict__ rec) {
  using R = ChRec<W, AT>;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
    int *ri = reinterpret_cast<int *>(rec + (size_t)t * R::bytes);
    AT *rd = reinterpret_cast<AT *>(rec + (size_t)t * R::bytes);
    const int4 d = desc[cmap[t]];
#pragma unroll
    for (int j = 0; j < W; ++j) ri[j] = code[(size_t)t * W + j];
    ri[R::row] = d.x;
    int k = 0;
    AT piv = AT(0);
    for (int e = d.y; e < d.z; ++e) {
      if (ja[e] == d.x)
        piv = a[e];
      else if (k < W)
        rd[R::val + k++] = a[e];
    }
    for (; k < W; ++k) rd[R::val + k] = AT(0);
    rd[R::piv] = piv;
  }
}

int *ch_progress() {
  static DevBufPtr buf;
  if (!buf) buf = dev_alloc(kChMax * sizeof(int));
  return buf && buf->p ? buf->as<int>() : nullptr;
}

int *sf_state() {  // [unused, error flag, unused, unused]
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

// ------------------------------------------------ chunked schedule (host)
//
// A partition of the rows into at most one chunk per CU is chosen among a few
// candidates by simulating the pipeline: every step of a chunk costs one unit
// and every wait on another chunk kHopCost more (a value crossing between
// workgroups goes through memory).  Candidates: contiguous ranges in solve
// order, and for matrices whose column offsets take few values (grids in
// lexicographic order) tiles over two of the offset strides, which keep the
// long dependency chains inside one chunk (e.g. the 256^3 grid's chains along
// z: 16 x 16 x 256 tiles cross 30 chunk boundaries on the critical path
// against 255 for z-slabs).
constexpr int kHopCost = 6;

// f(t0, t1) over contiguous ranges covering [0, n), one per host thread.
template <class F>
static void host_ranges(int n, F f) {
  const int T = std::max(1, std::min(host_threads(), n / 65536 + 1));
  std::vector<std::thread> th;
  for (int k = 1; k < T; ++k)
    th.emplace_back(f, (int)((long)n * k / T), (int)((long)n * (k + 1) / T));
  f(0, (int)((long)n / T));
  for (std::thread &t : th) t.join();
}

struct ChunkSched {
  int G = 0;
  std::vector<int> cbase, crow, cmap;  // rows in (chunk, level, row) order
  std::vector<int> cstep, sptr, slev, wptr, wchk, wlev;
  long long cost = -1;
};

// Steps and pruned waits of a partition; false when a chunk's step holds
// more than max_step rows.
static bool sched_build(ChunkSched &S, const std::vector<int> &chunk_of, int G,
                        const std::vector<int> &ia, const std::vector<int> &ja,
                        const std::vector<int> &level, const std::vector<int> &order,
                        int max_step) {
  const int n = (int)chunk_of.size();
  S.G = G;
  S.cbase.assign((size_t)G + 1, 0);
  for (int i = 0; i < n; ++i) S.cbase[(size_t)chunk_of[(size_t)i] + 1]++;
  for (int c = 0; c < G; ++c) S.cbase[(size_t)c + 1] += S.cbase[(size_t)c];
  S.crow.assign((size_t)n, 0);
  S.cmap.assign((size_t)n, 0);
  std::vector<int> fill(S.cbase.begin(), S.cbase.end() - 1);
  for (int t = 0; t < n; ++t) {  // order is (level, row): bucket stably by chunk
    const int i = order[(size_t)t];
    const int pos = fill[(size_t)chunk_of[(size_t)i]]++;
    S.crow[(size_t)pos] = i;
    S.cmap[(size_t)pos] = t;
  }
  S.cstep.assign((size_t)G + 1, 0);
  S.sptr.clear();
  S.slev.clear();
  S.wptr.clear();
  S.wchk.clear();
  S.wlev.clear();
  std::vector<int> running((size_t)G, -1), req((size_t)G, -1), touched, prod;
  for (int c = 0; c < G; ++c) {
    S.cstep[(size_t)c] = (int)S.slev.size();
    prod.clear();
    for (int t = S.cbase[(size_t)c]; t < S.cbase[(size_t)c + 1];) {
      const int lv = level[(size_t)S.crow[(size_t)t]];
      int u = t;
      touched.clear();
      while (u < S.cbase[(size_t)c + 1] && level[(size_t)S.crow[(size_t)u]] == lv) {
        const int i = S.crow[(size_t)u];
        for (int e = ia[i]; e < ia[i + 1]; ++e) {
          const int d = ja[e];
          if (d == i) continue;
          const int cd = chunk_of[(size_t)d];
          if (cd == c) continue;
          if (req[(size_t)cd] < 0) touched.push_back(cd);
          req[(size_t)cd] = std::max(req[(size_t)cd], level[(size_t)d]);
        }
        ++u;
      }
      if (u - t > max_step) return false;
      S.sptr.push_back(t);
      S.slev.push_back(lv);
      S.wptr.push_back((int)S.wchk.size());
      std::sort(touched.begin(), touched.end());
      for (int cd : touched) {
        if (req[(size_t)cd] > running[(size_t)cd]) {
          if (running[(size_t)cd] < 0) prod.push_back(cd);
          running[(size_t)cd] = req[(size_t)cd];
          S.wchk.push_back(cd);
          S.wlev.push_back(req[(size_t)cd]);
        }
        req[(size_t)cd] = -1;
      }
      t = u;
    }
    for (int cd : prod) running[(size_t)cd] = -1;
  }
  const int nsteps = (int)S.slev.size();
  S.cstep[(size_t)G] = nsteps;
  S.sptr.push_back(n);
  S.wptr.push_back((int)S.wchk.size());
  for (int c = 0; c < G; ++c) {  // the kernel's LDS tables
    const int k0 = S.cstep[(size_t)c], k1 = S.cstep[(size_t)c + 1];
    if (k1 - k0 > kChSteps || S.wptr[(size_t)k1] - S.wptr[(size_t)k0] > kChWaits) return false;
  }
  // pipeline simulation, steps taken in level order (every wait refers to a
  // lower level, so its producer step is already timed)
  int nlev = 0;
  for (int v : S.slev) nlev = std::max(nlev, v + 1);
  std::vector<int> lcnt((size_t)nlev + 1, 0), byl((size_t)nsteps), owner((size_t)nsteps);
  for (int c = 0; c < G; ++c)
    for (int k = S.cstep[(size_t)c]; k < S.cstep[(size_t)c + 1]; ++k) owner[(size_t)k] = c;
  for (int k = 0; k < nsteps; ++k) lcnt[(size_t)S.slev[(size_t)k] + 1]++;
  for (int l = 0; l < nlev; ++l) lcnt[(size_t)l + 1] += lcnt[(size_t)l];
  for (int k = 0; k < nsteps; ++k) byl[(size_t)lcnt[(size_t)S.slev[(size_t)k]]++] = k;
  std::vector<long long> tdone((size_t)nsteps, 0);
  long long cost = 0;
  for (int k : byl) {
    const int c = owner[(size_t)k];
    long long t0 = k > S.cstep[(size_t)c] ? tdone[(size_t)k - 1] : 0;
    for (int q = S.wptr[(size_t)k]; q < S.wptr[(size_t)k + 1]; ++q) {
      const int cd = S.wchk[(size_t)q], L = S.wlev[(size_t)q];
      // the producer's first step at a level >= L
      const int *b0 = S.slev.data() + S.cstep[(size_t)cd], *b1 = S.slev.data() + S.cstep[(size_t)cd + 1];
      const int kk = (int)(std::lower_bound(b0, b1, L) - S.slev.data());
      t0 = std::max(t0, tdone[(size_t)kk] + kHopCost);
    }
    tdone[(size_t)k] = t0 + 1;
    cost = std::max(cost, t0 + 1);
  }
  S.cost = cost;
  return true;
}

// The chunked schedule of a triangle whose rows hold at most 8 entries.
static void tri_chunks_build(TriPlan &p, const std::vector<int> &ia, const std::vector<int> &ja,
                             const std::vector<int> &level, const std::vector<int> &order) {
  const int n = p.n;
  if (n <= 0 || p.max_row > 5 || !p.has_diag) return;
  const int Gmax = std::min({device_cus(), kChMax, std::max(1, n / kChMinRows)});
  const bool lower = p.lower;
  auto pos_of = [&](int i) { return lower ? i : n - 1 - i; };
  std::vector<std::pair<std::vector<int>, int>> cands;
  {  // contiguous ranges in solve order
    int C = (n + Gmax - 1) / Gmax;
    C = (C + 63) / 64 * 64;
    const int G = (n + C - 1) / C;
    std::vector<int> co((size_t)n);
    for (int i = 0; i < n; ++i) co[(size_t)i] = pos_of(i) / C;
    cands.emplace_back(std::move(co), G);
  }
  {  // tiles over two offset strides
    std::vector<int> offs;
    for (int i = 0; i < n && offs.size() <= 16; ++i)
      for (int e = ia[i]; e < ia[i + 1]; ++e) {
        const int d = ja[e];
        if (d == i) continue;
        const int o = std::abs(i - d);
        if (std::find(offs.begin(), offs.end(), o) == offs.end()) offs.push_back(o);
      }
    if (offs.size() <= 16) {
      std::sort(offs.begin(), offs.end());
      for (size_t x = 0; x < offs.size(); ++x)
        for (size_t y = x + 1; y < offs.size(); ++y) {
          const int Oa = offs[x], Ob = offs[y];
          if (Oa < 2 || Ob % Oa || n % Ob) continue;
          const int nb = Ob / Oa;  // extent of the second coordinate
          for (int nI = 1; nI <= Gmax; nI *= 2) {
            const int nJ = Gmax / nI;
            if (nI > Oa || nJ > nb || nI * nJ < 2) continue;
            const int Ti = (Oa + nI - 1) / nI, Tj = (nb + nJ - 1) / nJ;
            const int gi = (Oa + Ti - 1) / Ti, gj = (nb + Tj - 1) / Tj;
            std::vector<int> co((size_t)n);
            for (int r = 0; r < n; ++r) {
              const int ps = pos_of(r);
              co[(size_t)r] = ((ps / Oa) % nb / Tj) * gi + (ps % Oa) / Ti;
            }
            cands.emplace_back(std::move(co), gi * gj);
          }
        }
    }
  }
  // the candidates are independent: one host thread each (the lowest cost
  // wins, ties to the earlier candidate, so the choice does not depend on
  // thread timing)
  std::vector<ChunkSched> built(cands.size());
  {
    std::atomic<size_t> next{0};
    auto work = [&]() {
      for (size_t ci; (ci = next.fetch_add(1)) < cands.size();)
        if (!sched_build(built[ci], cands[ci].first, cands[ci].second, ia, ja, level, order,
                         kChStepRows))
          built[ci] = ChunkSched();
    };
    const int T = std::max(1, std::min(host_threads(), (int)cands.size()));
    std::vector<std::thread> th;
    for (int t = 1; t < T; ++t) th.emplace_back(work);
    work();
    for (std::thread &t : th) t.join();
  }
  size_t best_i = 0;
  for (size_t ci = 0; ci < cands.size(); ++ci)
    if (built[ci].cost >= 0 && (built[best_i].cost < 0 || built[ci].cost < built[best_i].cost))
      best_i = ci;
  if (cands.empty() || built[best_i].cost < 0) return;
  ChunkSched best = std::move(built[best_i]);
  built.clear();
  if (getenv("LAP_SCHED_LOG"))
    fprintf(stderr, "[lapgfx] %s triangle n=%d: chunked schedule %zu of %zu, %d chunks, %zu steps, "
            "%zu waits, simulated length %lld (one launch per level: %d)\n",
            lower ? "lower" : "upper", n, best_i, cands.size(), best.G, best.slev.size(),
            best.wchk.size(), best.cost, p.nlevels);
  const std::vector<int> &chunk_of = cands[best_i].first;
  const int G = best.G;
  // dependency codes of the off-diagonal entries, records in schedule order:
  // -1 padding, >= 0 a row read from y, <= -2 the LDS ring slot -(code + 2)
  // of a row of this chunk solved fewer than kChRing steps earlier
  const int W = 4;  // off-diagonal entries per record
  std::vector<int> lstep((size_t)n), lpos((size_t)n);  // local step and position of each row
  for (int c = 0; c < G; ++c)
    for (int k = best.cstep[(size_t)c]; k < best.cstep[(size_t)c + 1]; ++k)
      for (int t = best.sptr[(size_t)k]; t < best.sptr[(size_t)k + 1]; ++t) {
        lstep[(size_t)best.crow[(size_t)t]] = k - best.cstep[(size_t)c];
        lpos[(size_t)best.crow[(size_t)t]] = t - best.sptr[(size_t)k];
      }
  std::vector<int> code((size_t)n * W, -1);
  host_ranges(n, [&](int t0, int t1) {  // records are independent: contiguous ranges per thread
    for (int t = t0; t < t1; ++t) {
      const int i = best.crow[(size_t)t];
      int k = 0;
      for (int e = ia[i]; e < ia[i + 1]; ++e) {
        const int d = ja[e];
        if (d == i) continue;  // the pivot is kept apart
        int cd = d;
        if (chunk_of[(size_t)d] == chunk_of[(size_t)i] &&
            lstep[(size_t)i] - lstep[(size_t)d] < kChRing)
          cd = -(2 + (lstep[(size_t)d] % kChRing) * kChStepRows + lpos[(size_t)d]);
        code[(size_t)t * W + k++] = cd;
      }
    }
  });
  // the progress counters and the error flag, allocated here (never inside
  // a graph capture of the solve)
  if (!ch_progress() || !sf_state()) return;
  if (best.wchk.empty()) {  // keep the device arrays non-empty
    best.wchk.push_back(0);
    best.wlev.push_back(-1);
  }
  TriPlan::Chunks &ch = p.ch;
  ch.nchunks = G;
  ch.nsteps = (int)best.slev.size();
  ch.nwaits = best.wptr.back();
  ch.w = W;
  ch.cost = best.cost;
  ch.cstep = dev_upload(best.cstep);
  ch.sptr = dev_upload(best.sptr);
  ch.slev = dev_upload(best.slev);
  ch.wptr = dev_upload(best.wptr);
  ch.wchk = dev_upload(best.wchk);
  ch.wlev = dev_upload(best.wlev);
  ch.crow = dev_upload(best.crow);
  ch.cmap = dev_upload(best.cmap);
  ch.code = dev_upload(code);
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
  std::vector<int> desc((size_t)n * 4);
  for (int t = 0; t < n; ++t) {
    const int i = order[(size_t)t];
    int piv = -1;  // the last diagonal entry, as the sequential loop keeps
    for (int e = ia[i]; e < ia[i + 1]; ++e)
      if (ja[e] == i) piv = e;
    desc[(size_t)t * 4] = i;
    desc[(size_t)t * 4 + 1] = ia[i];
    desc[(size_t)t * 4 + 2] = ia[i + 1];
    desc[(size_t)t * 4 + 3] = piv;
  }
  p->desc = dev_upload(desc);
  // segments of the long rows (tree-summed level launches)
  std::vector<int> segd;
  p->level_seg.assign((size_t)nlev + 1, 0);
  for (int l = 0; l < nlev; ++l) {
    for (int t = p->level_long[l]; t < p->level_ptr[l + 1]; ++t) {
      const int i = order[(size_t)t];
      const int ns = (ia[i + 1] - ia[i] + kTriSeg - 1) / kTriSeg;
      const int slot = (int)(segd.size() / 4);
      for (int g = 0; g < ns; ++g)
        segd.insert(segd.end(), {t - p->level_ptr[l], g, ns, slot});
    }
    p->level_seg[(size_t)l + 1] = (int)(segd.size() / 4);
  }
  const size_t nseg = segd.size() / 4;
  p->segd = dev_upload(segd);
  p->seg_part = dev_alloc(sizeof(double) * (nseg + 1));
  p->seg_tick = dev_upload(std::vector<int>(nseg + 1, 0));
  tri_chunks_build(*p, ia, ja, level, order);
  return p;
}

// The analysis of a device pattern (not cached).
static std::unique_ptr<TriPlan> tri_plan_make(int n, int nnz, const int *ia, const int *ja,
                                              bool lower) {
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
  return tri_analyse_host(n, h_ia, h_ja, lower);
}

const TriPlan *tri_plan(int n, int nnz, const int *ia, const int *ja, bool lower) {
  Key key{ia, ja, n, nnz, lower};
  auto it = g_plans.find(key);
  if (it != g_plans.end()) return it->second.get();
  auto plan = tri_plan_make(n, nnz, ia, ja, lower);
  if (!plan) return nullptr;
  TriPlan *raw = plan.get();
  g_plans[key] = std::move(plan);
  return raw;
}

std::pair<std::unique_ptr<TriPlan>, std::unique_ptr<TriPlan>> tri_analyse_pair(
    int n, const std::vector<int> &lia, const std::vector<int> &lja,
    const std::vector<int> &uia, const std::vector<int> &uja) {
  // the shared device buffers are allocated lazily: before the threads
  ch_progress();
  sf_state();
  std::unique_ptr<TriPlan> up;
  std::thread tu([&] { up = tri_analyse_host(n, uia, uja, false); });
  auto lp = tri_analyse_host(n, lia, lja, true);
  tu.join();
  return {std::move(lp), std::move(up)};
}

void tri_plan_pair(int n, int nnzL, const int *lia, const int *lja, int nnzU, const int *uia,
                   const int *uja) {
  const Key kl{lia, lja, n, nnzL, true}, ku{uia, uja, n, nnzU, false};
  const bool need_l = !g_plans.count(kl), need_u = !g_plans.count(ku);
  if (!need_l || !need_u) {  // one (or none) to build
    tri_plan(n, nnzL, lia, lja, true);
    tri_plan(n, nnzU, uia, uja, false);
    return;
  }
  ch_progress();
  sf_state();
  std::unique_ptr<TriPlan> up;
  std::thread tu([&] { up = tri_plan_make(n, nnzU, uia, uja, false); });
  auto lp = tri_plan_make(n, nnzL, lia, lja, true);
  tu.join();
  if (lp) g_plans[kl] = std::move(lp);
  if (up) g_plans[ku] = std::move(up);
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
  if (p.ch.nchunks > 0 && p.ch.w == 4) {  // the chunked schedule's records for these values
    // kBlock records of padding: the prefetch copies whole slots
    e.crec = dev_alloc((size_t)ChRec<4, double>::bytes * ((size_t)p.n + kBlock));
    e.xc = dev_alloc(sizeof(double) * ((size_t)p.n + kBlock));
    if (!e.crec || !e.crec->p || !e.xc || !e.xc->p) return;
    k_ch_pack<4, double><<<g, kBlock, 0, stream()>>>(p.n, p.ch.cmap->as<int>(),
                                                     p.desc->as<int4>(), ja, a,
                                                     p.ch.code->as<int>(), e.crec->as<char>());
  }
}

void tri_ch_pack32(const TriPlan &p, const int *ja, const float *a, TriChunks32 &c) {
  c.ok = false;
  if (p.n <= 0 || p.ch.nchunks <= 0 || p.ch.w != 4) return;
  const size_t bytes = (size_t)ChRec<4, float>::bytes * ((size_t)p.n + kBlock);
  if (!c.rec || c.rec->bytes < bytes) c.rec = dev_alloc(bytes);
  const size_t xb = sizeof(float) * ((size_t)p.n + kBlock);
  if (!c.xc || c.xc->bytes < xb) c.xc = dev_alloc(xb);
  if (!c.rec || !c.rec->p || !c.xc || !c.xc->p) return;
  const unsigned g = grid_for((size_t)p.n, 1, kBlock, 65536);
  k_ch_pack<4, float><<<g, kBlock, 0, stream()>>>(p.n, p.ch.cmap->as<int>(), p.desc->as<int4>(),
                                                  ja, a, p.ch.code->as<int>(),
                                                  c.rec->as<char>());
  c.ok = true;
}

// One launch per level: a lane per short row, then a workgroup per long row
// (sequential order) or per long-row segment (option sptrsv_long 1).
template <class AT, class VT>
static void tri_level_launch(const TriPlan &p, const int *ja, const AT *a, const double *diag,
                             const VT *x, VT *y, const int *done) {
  const int4 *desc = p.desc->as<int4>();
  const bool tree = options().sptrsv_long == 1;
  const int4 *segd = p.segd->as<int4>();
  double *part = p.seg_part->as<double>();
  int *tick = p.seg_tick->as<int>();
  for (int l = 0; l < p.nlevels; ++l) {
    const int cnt = p.level_long[l] - p.level_ptr[l];
    const int nlong = tree ? p.level_seg[(size_t)l + 1] - p.level_seg[(size_t)l]
                           : p.level_ptr[l + 1] - p.level_long[l];
    const unsigned g = (unsigned)((cnt + kBlock - 1) / kBlock + nlong);
    if (tree)
      k_tri_level<AT, VT, true><<<g, kBlock, 0, stream()>>>(cnt, nlong, desc + p.level_ptr[l],
                                                            ja, a, diag, x, y, done,
                                                            segd + p.level_seg[(size_t)l], part,
                                                            tick);
    else
      k_tri_level<AT, VT><<<g, kBlock, 0, stream()>>>(cnt, nlong, desc + p.level_ptr[l], ja, a,
                                                      diag, x, y, done, nullptr, nullptr,
                                                      nullptr);
  }
}

// Which path tri_solve_launch takes.
enum TriPath { TRI_CHUNKS, TRI_LEVELS_ELL, TRI_LEVELS };

static TriPath tri_path(const TriPlan &p, const TriEll *ell, bool in_place) {
  const int mode = options().sptrsv_mode;
  if (mode == 3 && ell && ell->w && ell->crec && ell->xc && p.ch.nchunks > 0 && p.ch.w == 4 &&
      sf_state() && ch_progress())
    return TRI_CHUNKS;
  (void)in_place;  // the chunked kernel reads x from its own gathered copy
  if (ell && ell->w) return TRI_LEVELS_ELL;
  return TRI_LEVELS;
}

int tri_launches(const TriPlan &p, const TriEll *ell, bool in_place) {
  switch (tri_path(p, ell, in_place)) {
    case TRI_CHUNKS: return 3;
    default: return p.nlevels;
  }
}

void tri_solve_launch(const TriPlan &p, const int *ia, const int *ja, const double *a,
                      const double *diag, const double *x, double *y, const int *done,
                      const TriEll *ell) {
  (void)ia;  // the plan's descriptors carry every row's range
  const int *order = p.order->as<int>();
  const TriPath path = tri_path(p, ell, x == y);
  if (path == TRI_CHUNKS) {
    int *st = sf_state(), *prog = ch_progress();
    {
      const TriPlan::Chunks &ch = p.ch;
      k_ch_reset<<<(ch.nchunks + kBlock - 1) / kBlock, kBlock, 0, stream()>>>(ch.nchunks, prog,
                                                                             done);
      k_ch_gather<double><<<2048, kBlock, 0, stream()>>>(
          p.n, ch.crow->as<int>(), x, ell->xc->as<double>(), done);
#define LAP_CH_ARGS                                                                            \
  ch.cstep->as<int>(), ch.sptr->as<int>(), ch.slev->as<int>(), ch.wptr->as<int>(),             \
      ch.wchk->as<int>(), ch.wlev->as<int>(), ell->crec->as<char>(),                           \
      ell->xc->as<double>(), y, prog, st, done
      k_tri_chunks<double, double, 4, 3, 4><<<ch.nchunks, kBlock + kWave, 0, stream()>>>(
          LAP_CH_ARGS);
#undef LAP_CH_ARGS
      return;
    }
  }
  if (path == TRI_LEVELS_ELL) {
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
  tri_level_launch<double, double>(p, ja, a, diag, x, y, done);
}

template <class VT>
static void tri_mixed_launch(const TriPlan &p, const int *ja, const float *a, const VT *x, VT *y,
                             const int *done) {
  tri_level_launch<float, VT>(p, ja, a, nullptr, x, y, done);
}

static bool tri_chunks32_ready(const TriPlan &p, const TriChunks32 *c) {
  return c && c->ok && options().sptrsv_mode == 3 && p.ch.nchunks > 0 && p.ch.w == 4 &&
         sf_state() && ch_progress();
}

int tri_launches32(const TriPlan &p, const TriChunks32 *c) {
  return tri_chunks32_ready(p, c) ? 3 : p.nlevels;
}

void tri_solve32_launch(const TriPlan &p, const int *ja, const float *a, const float *x, float *y,
                        const int *done, const TriChunks32 *c) {
  if (tri_chunks32_ready(p, c)) {
    const TriPlan::Chunks &ch = p.ch;
    int *st = sf_state(), *prog = ch_progress();
    k_ch_reset<<<(ch.nchunks + kBlock - 1) / kBlock, kBlock, 0, stream()>>>(ch.nchunks, prog,
                                                                           done);
    k_ch_gather<float><<<2048, kBlock, 0, stream()>>>(p.n, ch.crow->as<int>(), x,
                                                      c->xc->as<float>(), done);
    k_tri_chunks<float, float, 4, 3, 4><<<ch.nchunks, kBlock + kWave, 0, stream()>>>(
        ch.cstep->as<int>(), ch.sptr->as<int>(), ch.slev->as<int>(), ch.wptr->as<int>(),
        ch.wchk->as<int>(), ch.wlev->as<int>(), c->rec->as<char>(), c->xc->as<float>(), y, prog,
        st, done);
    return;
  }
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
  tri_plan_pair(n, nnzL, lia, lja, nnzU, uia, uja);
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
