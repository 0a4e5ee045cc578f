// lap_core.hpp -- internal definitions shared by the gfx950 backend.
//
// Execution model: one HIP stream per process (a blocking stream at default
// priority, so it orders with work the caller put on the legacy default
// stream), 64-lane wavefronts, 256-thread workgroups unless stated.  Every
// public entry point leaves the stream synchronised on return.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <initializer_list>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/lap_gfx950.h"

namespace lap {

constexpr int kWave = 64;
constexpr int kBlock = 256;

// ------------------------------------------------------------------ errors
void set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
bool hip_ok(hipError_t e, const char *what, const char *file, int line);
#define LAP_HIP(expr) ::lap::hip_ok((expr), #expr, __FILE__, __LINE__)

// ------------------------------------------------------------------ runtime
hipStream_t stream();
int host_threads();  // host threads the library's parallel host code uses
int device_cus();
void sync();  // hipStreamSynchronize(stream()) with error capture
// True when p is device (or managed) memory of the HIP runtime.  Used to
// refuse host pointers before any kernel could dereference them.
bool is_device_ptr(const void *p);
// Records an error and returns false unless every listed pointer is device
// memory with at least `bytes` bytes from p to the end of its allocation;
// every public entry point checks its caller pointers with it before
// launching anything, so a host pointer or a short buffer is refused, never
// read by a kernel.
struct PtrArg {
  const void *p;
  const char *name;
  size_t bytes = 0;  // required extent from p (0: not checked)
};
bool check_device(const char *fn, std::initializer_list<PtrArg> args);

// Device buffer owned by the library (RAII over hipMalloc/hipFree).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b);
  ~DevBuf();
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  template <class T>
  T *as() const { return static_cast<T *>(p); }
  void upload(const void *src, size_t b, size_t off = 0);
};
using DevBufPtr = std::shared_ptr<DevBuf>;
DevBufPtr dev_alloc(size_t bytes);
template <class T>
DevBufPtr dev_upload(const std::vector<T> &v) {
  auto b = dev_alloc(v.size() * sizeof(T) + 16);
  if (!v.empty()) b->upload(v.data(), v.size() * sizeof(T));
  return b;
}
template <class T>
std::vector<T> dev_download(const T *dptr, size_t count) {
  std::vector<T> h(count);
  if (count) {
    LAP_HIP(hipMemcpyAsync(h.data(), dptr, count * sizeof(T), hipMemcpyDeviceToHost, stream()));
    sync();
  }
  return h;
}

// Two pinned host ints for polling device flags (allocated once).
int *pinned_flags();

// Cached n-vectors reused across solves (avoids hipMalloc in timed regions).
double *scratch_vec(const char *tag, size_t n);

// ------------------------------------------------------------ reductions
// Fence-free deterministic reductions.  A producer kernel writes one partial
// sum per workgroup with a plain store; the consumer kernel (the next launch
// on the stream, so the kernel boundary makes the partials visible) sums the
// partials in a fixed order.  When the consumer is a full grid, every
// workgroup performs the same sum redundantly and gets the same bits.  This
// avoids agent-scope fences inside kernels, which on gfx950 write back and
// invalidate the XCD's L2 (buffer_wbl2 / buffer_inv) once per workgroup.
constexpr int kMaxPartialBlocks = 8192;
struct ReduceWs {
  double *partials;  // [kMaxPartialBlocks * 4]
  double *scalars;   // [64] general device scalars
};
ReduceWs &reduce_ws();
// The current device's address of lap_dot_noalloc's static scratch array
// (looked up once per device; null and an error recorded on failure).
double *dot_scratch();

// -------------------------------------------------------- device helpers
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Streamed vector elements: n-long vectors that a kernel reads or writes
// once, far larger than the L2s.  Non-temporal loads and stores keep them
// from displacing lines that are reused (the SpMV's gathers of p).  Values
// are unchanged; only the cache policy differs.
template <class T>
__device__ __forceinline__ T ld_nt(const T *p) { return __builtin_nontemporal_load(p); }
template <class T>
__device__ __forceinline__ void st_nt(T *p, T v) { __builtin_nontemporal_store(v, p); }

// s (+ or -) p[k0] (+ or -) p[k0 + 1] ... (+ or -) p[k1 - 1], one term at a
// time in that order: the sequential loop's result bit for bit.  The loads
// (p is LDS) are issued eight at a time ahead of the dependent adds, so the
// chain pays the add latency rather than a load round trip per term.
template <bool SUB = false>
__device__ __forceinline__ double lds_seq_sum(double s, const double *p, int k0, int k1) {
  int k = k0;
  for (; k + 8 <= k1; k += 8) {
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = p[k + j];
#pragma unroll
    for (int j = 0; j < 8; ++j) s = SUB ? s - v[j] : s + v[j];
  }
  if (k + 4 <= k1) {
    double v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = p[k + j];
#pragma unroll
    for (int j = 0; j < 4; ++j) s = SUB ? s - v[j] : s + v[j];
    k += 4;
  }
  for (; k < k1; ++k) s = SUB ? s - p[k] : s + p[k];  // 0..3 left: short rows
  return s;
}

// Sum over a block of BS threads; the result is valid in thread 0.
// Fixed summation order, so deterministic.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double *lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < BS / kWave; ++i) r += lds[i];
  }
  __syncthreads();
  return r;
}

// Sum of partials[0..cnt) (stride `stride` doubles) in a fixed order,
// returned to every thread of the block.
template <int BS>
__device__ __forceinline__ double sum_partials(const double *partials, int cnt, int stride,
                                               double *lds) {
  double s = 0.0;
  for (int i = threadIdx.x; i < cnt; i += BS) s += partials[(size_t)i * stride];
  s = block_sum<BS>(s, lds);
  if (threadIdx.x == 0) lds[0] = s;
  __syncthreads();
  const double r = lds[0];
  __syncthreads();
  return r;
}

// Work distribution of a persistent grid over nblk logical blocks (every
// logical block is visited exactly once).  Workgroups are dispatched
// round-robin over the 8 XCDs (XCD = id mod 8), each XCD with its own L2.
//  mode 0 cyclic:   workgroup w takes blocks w, w+G, w+2G, ...
//  mode 1 per-XCD:  the workgroups of XCD k walk the k-th contiguous eighth.
//  mode 2 blocked:  workgroup w takes the contiguous chunk
//                   [w*C, (w+1)*C), C = ceil(nblk/G).  On a grid-ordered
//                   matrix the rows that share x entries (neighbouring lines
//                   and planes) then sit in chunks of workgroups on one XCD.
//  modes 3/4/5 XCD groups: blocks are dealt to the XCDs in groups of g
//                   consecutive blocks (g = 4, 8, 2), group q to XCD q mod 8,
//                   and each XCD's workgroups walk their groups cyclically.
//                   Neighbouring blocks mostly share an XCD, and blocks
//                   8*g*m apart always do, while every XCD still streams
//                   addresses spread over the whole matrix.
// Loops run s over [begin, end) with stride step and process block(s),
// skipping block(s) >= nblk (only the last group can overrun).
struct BlockWalk {
  int begin, end, step;
  int nblk;     // row blocks of the plan
  int g = 0;    // group size of walk 3 (0: s is the block)
  int gk = 0;   // first block of this XCD's first group
  __device__ __forceinline__ int block(int s) const {
    return g ? (s / g) * (8 * g) + gk + (s % g) : s;
  }
};
__device__ __forceinline__ BlockWalk block_walk(int nblk, int mode) {
  const int g = gridDim.x, id = blockIdx.x;
  if (mode == 3 && g >= 8) {
    const int gs = 4;
    const int k = id % 8, lw = id / 8;
    const int nw = g / 8 + (k < g % 8 ? 1 : 0);
    const int base = gs * k;
    const int groups = base < nblk ? (nblk - base + 8 * gs - 1) / (8 * gs) : 0;
    BlockWalk w{lw, groups * gs, nw, nblk};
    w.g = gs;
    w.gk = base;
    return w;
  }
  return BlockWalk{id, nblk, g, nblk};
}

// -------------------------------------------------------------- options
struct Options {
  int spmv_rows = 256;   // rows per SpMV block (<= kBlock * 2)
  int spmv_walk = 0;     // row-block walk of the persistent SpMV grid: 0 cyclic,
                         // 3 groups of 4 consecutive blocks per XCD
  int vec_grid = 768;    // workgroups of the streaming vector kernels (3 per CU)
  int spmv_wg_per_cu = 8;  // persistent SpMV grid: workgroups per CU
  int graph_iters = 16;  // PCG iterations per graph launch
  int use_graph = 1;
  int spmm_wg_per_cu = 4;  // block CG SpMM grid: workgroups per CU
  int sptrsv_mode = 3;     // triangular solves: 0 one launch per level, 3 one
                           // persistent launch over chunks of rows with
                           // point-to-point waits
  int fuse_sweeps = 1;     // it_jacobi / GS_it / GS_it2: one launch per sweep
  int mixed_vectors = 0;   // lap_pcg_mixed inner vectors: 0 FP32, 1 FP64
  int sptrsv_long = 0;     // level solves' rows > kTriLongRow entries: 0 summed
                           // in CSR order (bitwise), 1 as a workgroup tree
  int spmv_format = 0;     // PCG SpMV: 0 CSR values, 1 value-indexed when the
                           // matrix has at most 255 distinct values, 2 also
                           // the columns as offset indices (<= 255 distinct
                           // col - row, rows <= 64 entries), else as 1
};
Options &options();

inline unsigned grid_for(size_t n, int per_thread = 4, int bs = kBlock, unsigned cap = 0) {
  size_t g = (n + (size_t)bs * per_thread - 1) / ((size_t)bs * per_thread);
  if (g < 1) g = 1;
  if (cap && g > cap) g = cap;
  return (unsigned)g;
}

// out[0..n] = exclusive scan of in[0..n) on the device (out[n] = total);
// returns the total, or -1 after a HIP error.  ws: scan_ws_bytes(n) bytes.
size_t scan_ws_bytes(long long n);
long long scan_exclusive(const int *in, long long n, int *out, long long *ws);

// Drops every cached plan (SpMV, triangular, IC0, mixed, batch) whose key
// holds the device address p.
void forget_plans(const void *p);

// ------------------------------------------------------------- SpMV plan
struct SpmvPlan {
  int n = 0, nnz = 0;
  int nblocks = 0;        // row blocks
  DevBufPtr rowblk;       // int[nblocks + 1]: first row of each block
  int max_row = 0;
};
// Finds or builds the plan for (ia, ja, n, nnz) (device arrays).
const SpmvPlan *spmv_plan(int n, int nnz, const int *ia, const int *ja);

// y = alpha*A*x + beta*y on the library stream (no sync).
void spmv_launch(const SpmvPlan &p, const int *ia, const int *ja, const double *a,
                 const double *x, double *y, double alpha, double beta,
                 const int *done = nullptr);
// Persistent SpMV grid: min(row blocks, 8 workgroups per CU).
unsigned spmv_grid(const SpmvPlan &p);

// ------------------------------------------------- value-indexed CSR
// Encodes a[0..nnz) as one-byte indices into a table of its distinct values
// (at most 255, bit patterns ascending).  Returns the table size, 0 when the
// values do not fit, -1 after an error.
int vi_encode(long long nnz, const double *a, unsigned char *vi, double *dict);
// Delta-indexed CSR: one 16-bit code per nonzero, low byte = index into
// vdict (the distinct values, as vi_encode), high byte = index into ddict
// (the distinct column offsets col - row, ascending as unsigned).  Returns 1
// when encoded, 0 when either table would exceed 255 entries or a row is
// longer than max_row, -1 after an error.
int dvi_encode(int n, long long nnz, const int *ia, const int *ja, const double *a, int max_row,
               unsigned short *code, double *vdict, int *ddict, int *nv, int *ndl);

// --------------------------------------------------- SpMV-only sweeps
// Row blocks over a pair of CSR matrices (the strict L and U of a split),
// at most kSpmvTile nonzeros of both per block (a longer row gets a block of
// its own).
struct PairPlan {
  int n = 0;
  int nblocks = 0;
  DevBufPtr rowblk;  // int[nblocks + 1]: first row of each block
};
const PairPlan *pair_plan(int n, int nl, const int *lia, int nu, const int *uia);
// y_i = base_i - sum_j a_ij x_j (CSR order), stored per `mode`:
// 0 out = y * d_r; 1 out = y; 2 out = y, out2 = y * d_r;
// 3 out2 = (first ? 0 : out2) + y * d_r.
void sweep_one_launch(const SpmvPlan &p, const int *ia, const int *ja, const double *a,
                      const double *x, const double *base, const double *d_r, double *out,
                      double *out2, int mode, bool first, const int *done);
// out_i = (base_i - (sum_U a_ij x_j + sum_L a_ij x_j)) * d_r_i
void sweep_pair_launch(const PairPlan &p, const int *lia, const int *lja, const double *la,
                       const int *uia, const int *uja, const double *ua, const double *x,
                       const double *base, const double *d_r, double *out, const int *done);

// --------------------------------------------------- triangular solve plan
struct TriPlan {
  int n = 0, nnz = 0;
  bool lower = true;
  int nlevels = 0;
  std::vector<int> level_ptr;  // host, nlevels + 1
  DevBufPtr desc;              // int4[n] in level order: {row, ia[row], ia[row+1], 0}
  int max_row = 0;             // longest row (entries, pivot included)
  DevBufPtr order;             // int[n]: rows sorted by level
  int max_level_rows = 0;
  // Within each level, rows of at most kTriLongRow entries come first, then
  // the longer ones: level l's long rows are order[level_long[l] ..
  // level_ptr[l + 1]).  A long row is solved by a whole workgroup.
  std::vector<int> level_long;  // host, nlevels
  // Long rows split into segments of kTriSeg entries for the tree-summed
  // level launches (option sptrsv_long 1): level l's segments are
  // segd[level_seg[l] .. level_seg[l + 1]), int4 {position in the level,
  // segment, segments of the row, slot}; a row's segments are consecutive
  // and its slot is the index of its first one, which addresses the row's
  // partial sums (seg_part[slot + segment], double) and ticket
  // (seg_tick[slot], int, zero between solves).
  std::vector<int> level_seg;  // host, nlevels + 1
  DevBufPtr segd, seg_part, seg_tick;
  bool has_diag = true;        // every row holds its pivot
  // Chunked schedule (one persistent workgroup per chunk of consecutive rows
  // in solve order; built when every row holds at most 8 entries).  Chunk c
  // runs its rows level by level ("steps"); before a step it waits until
  // each earlier chunk it depends on has published the level it needs.
  struct Chunks {
    int nchunks = 0, nsteps = 0, nwaits = 0;
    DevBufPtr cstep;  // int[nchunks + 1]: first step of each chunk
    DevBufPtr sptr;   // int[nsteps + 1]: first position of each step
    DevBufPtr slev;   // int[nsteps]: the step's level
    DevBufPtr wptr;   // int[nsteps + 1]: first wait of each step
    DevBufPtr wchk;   // int[nwaits]: chunk waited on
    DevBufPtr wlev;   // int[nwaits]: level it must have published
    DevBufPtr crow;   // int[n]: rows in (chunk, level, row) order
    DevBufPtr cmap;   // int[n]: level-order position (into desc) of crow[t]
    DevBufPtr code;   // int[n * w]: each record's dependencies (see sptrsv.hip)
    int w = 0;        // record width
    long long cost = 0;  // simulated pipeline length (steps)
  } ch;
};
// Rows longer than this many entries (pivot included) are solved by a
// workgroup each in the level launches.
constexpr int kTriLongRow = 32;
// Entries per workgroup of a tree-summed long row (a multiple of kBlock).
constexpr int kTriSeg = 1024;
const TriPlan *tri_plan(int n, int nnz, const int *ia, const int *ja, bool lower);
// Both plans of a lower/upper pair (GS_std's splitting), analysed at once.
void tri_plan_pair(int n, int nnzL, const int *lia, const int *lja, int nnzU, const int *uia,
                   const int *uja);
// A triangle's rows in level order as fixed-width records (w = 4 or 8
// entries, padded with column -1): ja[t*w + k], a[t*w + k] for the row at
// level-order position t.  Built for one set of values; w == 0: not built.
struct TriEll {
  int w = 0;
  DevBufPtr ja, a;
  DevBufPtr crec;  // the chunked schedule's records for these values
  DevBufPtr xc;    // a solve's right-hand side in that schedule's order
};
// Packs (ja, a) of the plan's matrix when its rows hold at most 8 entries.
void tri_ell_build(const TriPlan &p, const int *ja, const double *a, TriEll &e);
// y = T^{-1} x (level-scheduled); ctl_done (device int*, may be null) makes
// every launch a no-op once *ctl_done != 0.
// `ell` (may be null): the packed rows of the same values, used by the
// level launches when built.
void tri_solve_launch(const TriPlan &p, const int *ia, const int *ja, const double *a,
                      const double *diag, const double *x, double *y, const int *ctl_done,
                      const TriEll *ell = nullptr);
// The number of kernel launches tri_solve_launch makes for these arguments
// (x == y: in place).
int tri_launches(const TriPlan &p, const TriEll *ell, bool in_place);
// The same level-scheduled solve over FP32 values and vectors with FP64 row
// sums: y = (float)((x - sum a*y) / pivot), pivot in the row.
// The chunked schedule's records over FP32 values (for tri_solve32_launch)
// and a buffer for the right-hand side in schedule order.
struct TriChunks32 {
  DevBufPtr rec, xc;
  bool ok = false;
};
void tri_ch_pack32(const TriPlan &p, const int *ja, const float *a, TriChunks32 &c);
// With `c` packed for these values and sptrsv_mode 3: one chunked launch
// (bitwise the level launches); otherwise the level launches.
void tri_solve32_launch(const TriPlan &p, const int *ja, const float *a, const float *x, float *y,
                        const int *done, const TriChunks32 *c = nullptr);
int tri_launches32(const TriPlan &p, const TriChunks32 *c);
// Sync-free mode: nonzero once a dependency wait ran out (the result is then
// invalid); reset by tri_clear_error.
int tri_error();
void tri_clear_error();

// -------------------------------------------------------------- IC0 plan
struct IcPlan {
  int n = 0;
  int nnzL = 0;
  DevBufPtr lia, lja, lval;  // L (diagonal last)
  DevBufPtr uia, uja, uval;  // U = L^T (diagonal first)
  const TriPlan *lplan = nullptr, *uplan = nullptr;
  TriEll lell, uell;  // the factors' rows packed in level order
  bool ok = true;
};
const IcPlan *ic_plan_find(const double *vals);
const IcPlan *ic_factor(int n, int nnzA, const int *ia, const int *ja, double *a);
void ic_apply_launch(const IcPlan &p, const double *x, double *y, double *aux,
                     const int *ctl_done);

// ---------------------------------------------------- elementwise launches
// `done` (device int*, may be null): the launch is a no-op once *done != 0.
void launch_axpy(int n, double alpha, const double *x, double *y, const int *done = nullptr);
void launch_scal(int n, double alpha, double *v, const int *done = nullptr);
void launch_vec_vec(int n, const double *x, const double *y, double *res,
                    const int *done = nullptr);
void launch_copy(int n, const double *src, double *dst, const int *done = nullptr);
void launch_fill(int n, double v, double *dst, const int *done = nullptr);
// Deterministic dot; result left in *out (device).
void launch_dot(int n, const double *v, const double *w, double *out);

}  // namespace lap
