// vi.hip -- value-indexed CSR (CSR-VI): the matrix values replaced by a
// one-byte index into a table of the distinct values.
//
// A graph Laplacian of an unweighted graph has a handful of distinct values
// (-1 and the degrees plus the shift); the 256^3 grid Laplacian has 5.  CSR-VI
// (Kourtis, Goumas and Koziris, "Optimizing sparse matrix-vector
// multiplication using index and value compression", CF 2008) stores an
// 8-bit index per nonzero and the table once, so an SpMV streams 5 bytes per
// nonzero instead of 12.  The table holds the very doubles of the matrix, so
// every product and sum is the CSR one bit for bit.
//
// Encoding, on the device, once per solve (the values may change between
// solves at the same address):
//   1. every workgroup collects the distinct value bit patterns of its
//      share in an LDS hash set and merges them into a global hash set of
//      kViSlots slots (64-bit compare-and-swap); more than kViMax distinct
//      values (or the empty-slot pattern itself) raise the overflow flag;
//   2. the host sorts the distinct patterns (a fixed order) and uploads the
//      table;
//   3. every nonzero's index is found by binary search in the LDS copy of
//      the table.
// A matrix with more than kViMax distinct values is not encoded and the
// caller keeps the CSR values.
#include <algorithm>
#include <cstring>

#include "lap_core.hpp"

namespace lap {

namespace {

constexpr int kViMax = 255;
constexpr int kViSlots = 1024;   // global hash set
constexpr int kViLds = 512;      // per-workgroup hash set
constexpr unsigned long long kViEmpty = ~0ull;

__device__ __forceinline__ unsigned vi_hash(unsigned long long k, int bits) {
  return (unsigned)((k * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

// Inserts key k into the workgroup's LDS hash set; raises `over` when the
// set would exceed kViMax keys or k is the empty-slot pattern.
__device__ __forceinline__ void vi_insert(unsigned long long k, unsigned long long *set,
                                          int *count, int *over) {
  if (k == kViEmpty) {
    *over = 1;
    return;
  }
  unsigned h = vi_hash(k, 9);
  for (int probe = 0; probe < kViLds; ++probe, h = (h + 1) & (kViLds - 1)) {
    const unsigned long long cur = set[h];
    if (cur == k) return;
    if (cur == kViEmpty) {
      const unsigned long long prev = atomicCAS(&set[h], kViEmpty, k);
      if (prev == kViEmpty) {
        if (atomicAdd(count, 1) >= kViMax) *over = 1;
        return;
      }
      if (prev == k) return;
    }
  }
}

// Merges the workgroup's LDS set into the global set (after a barrier).
__device__ __forceinline__ void vi_merge(const unsigned long long *set, int over,
                                         unsigned long long *table, int *state) {
  if (over) {
    if (threadIdx.x == 0) atomicOr(&state[1], 1);
    return;
  }
  for (int i = threadIdx.x; i < kViLds; i += kBlock) {
    const unsigned long long k = set[i];
    if (k == kViEmpty) continue;
    unsigned h = vi_hash(k, 10);
    for (int probe = 0; probe < kViSlots; ++probe, h = (h + 1) & (kViSlots - 1)) {
      const unsigned long long prev = atomicCAS(&table[h], kViEmpty, k);
      if (prev == kViEmpty) {
        if (atomicAdd(&state[0], 1) >= kViMax) atomicOr(&state[1], 1);
        break;
      }
      if (prev == k) break;
    }
  }
}

// The column offsets (column - row) of a CSR matrix as keys (a 32-bit
// pattern, so never the empty-slot one); state as k_vi_collect's.
__global__ __launch_bounds__(kBlock) void k_delta_collect(int n, const int *__restrict__ ia,
                                                          const int *__restrict__ ja,
                                                          unsigned long long *table,
                                                          int *state) {
  __shared__ unsigned long long set[kViLds];
  __shared__ int count, over;
  for (int i = threadIdx.x; i < kViLds; i += kBlock) set[i] = kViEmpty;
  if (threadIdx.x == 0) {
    count = 0;
    over = 0;
  }
  __syncthreads();
  int longest = 0;
  for (int row = blockIdx.x * kBlock + threadIdx.x; row < n; row += gridDim.x * kBlock) {
    longest = max(longest, ia[row + 1] - ia[row]);
    for (int e = ia[row]; e < ia[row + 1] && !over; ++e)
      vi_insert((unsigned long long)(unsigned)(ja[e] - row), set, &count, &over);
    if (over) break;
  }
  atomicMax(&state[2], longest);
  __syncthreads();
  vi_merge(set, over, table, state);
}

// state: [0] distinct count, [1] overflow
__global__ __launch_bounds__(kBlock) void k_vi_collect(long long nnz, const double *__restrict__ a,
                                                       unsigned long long *table, int *state) {
  __shared__ unsigned long long set[kViLds];
  __shared__ int count, over;
  for (int i = threadIdx.x; i < kViLds; i += kBlock) set[i] = kViEmpty;
  if (threadIdx.x == 0) {
    count = 0;
    over = 0;
  }
  __syncthreads();
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < nnz; e += stride) {
    vi_insert((unsigned long long)__double_as_longlong(a[e]), set, &count, &over);
    if (over) break;
  }
  __syncthreads();
  vi_merge(set, over, table, state);
}

__global__ __launch_bounds__(kBlock) void k_vi_fill(unsigned long long *table, int *state) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < kViSlots; i += gridDim.x * kBlock)
    table[i] = kViEmpty;
  if (blockIdx.x == 0 && threadIdx.x < 3) state[threadIdx.x] = 0;
}

// position of key k in the ascending table d[0..nd) (k is present)
__device__ __forceinline__ int vi_find(const unsigned long long *d, int nd,
                                       unsigned long long k) {
  int lo = 0, hi = nd - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (d[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// code[e] = (index of a[e] in vkeys) | (index of ja[e] - row in dkeys) << 8
__global__ __launch_bounds__(kBlock) void k_dvi_encode(int n, const int *__restrict__ ia,
                                                       const int *__restrict__ ja,
                                                       const double *__restrict__ a,
                                                       const unsigned long long *__restrict__ vkeys,
                                                       int nv,
                                                       const unsigned long long *__restrict__ dkeys,
                                                       int ndl, unsigned short *__restrict__ code) {
  __shared__ unsigned long long vd[kViMax], dd[kViMax];
  for (int i = threadIdx.x; i < nv; i += kBlock) vd[i] = vkeys[i];
  for (int i = threadIdx.x; i < ndl; i += kBlock) dd[i] = dkeys[i];
  __syncthreads();
  for (int row = blockIdx.x * kBlock + threadIdx.x; row < n; row += gridDim.x * kBlock)
    for (int e = ia[row]; e < ia[row + 1]; ++e) {
      const int v = vi_find(vd, nv, (unsigned long long)__double_as_longlong(a[e]));
      const int d = vi_find(dd, ndl, (unsigned long long)(unsigned)(ja[e] - row));
      code[e] = (unsigned short)(v | (d << 8));
    }
}

__global__ __launch_bounds__(kBlock) void k_vi_encode(long long nnz, const double *__restrict__ a,
                                                      const unsigned long long *__restrict__ dict,
                                                      int nd, unsigned char *__restrict__ vi) {
  __shared__ unsigned long long d[kViMax];
  for (int i = threadIdx.x; i < nd; i += kBlock) d[i] = dict[i];
  __syncthreads();
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < nnz; e += stride) {
    const unsigned long long k = (unsigned long long)__double_as_longlong(a[e]);
    int lo = 0, hi = nd - 1;
    while (lo < hi) {  // every value is in the table
      const int mid = (lo + hi) >> 1;
      if (d[mid] < k)
        lo = mid + 1;
      else
        hi = mid;
    }
    vi[e] = (unsigned char)lo;
  }
}

}  // namespace

// Encodes a[0..nnz) into vi (nnz bytes) and dict (<= 255 doubles, bit
// patterns sorted ascending).  Returns the number of table entries, 0 when
// the values do not fit (more than 255 distinct), or -1 after an error.
int vi_encode(long long nnz, const double *a, unsigned char *vi, double *dict) {
  static DevBufPtr table, state;
  if (!table) {
    table = dev_alloc(sizeof(unsigned long long) * kViSlots);
    state = dev_alloc(3 * sizeof(int));
    if (!table || !state || !table->p || !state->p) return -1;
  }
  unsigned long long *t = table->as<unsigned long long>();
  int *st = state->as<int>();
  k_vi_fill<<<4, kBlock, 0, stream()>>>(t, st);
  if (nnz > 0) {
    const unsigned g = grid_for((size_t)nnz, 16, kBlock, 1024);
    k_vi_collect<<<g, kBlock, 0, stream()>>>(nnz, a, t, st);
  }
  std::vector<unsigned long long> h = dev_download(t, kViSlots);
  std::vector<int> hs = dev_download(st, 2);
  if (lap_last_status() == LAP_ERR_HIP) return -1;
  if (hs[1] || hs[0] > kViMax) return 0;
  std::vector<unsigned long long> keys;
  for (unsigned long long k : h)
    if (k != kViEmpty) keys.push_back(k);
  std::sort(keys.begin(), keys.end());
  if (keys.empty()) keys.push_back(0);  // nnz == 0: a harmless one-entry table
  if (!LAP_HIP(hipMemcpyAsync(dict, keys.data(), keys.size() * 8, hipMemcpyHostToDevice,
                              stream())))
    return -1;
  if (nnz > 0 && vi) {  // vi == nullptr: only the table is wanted
    const unsigned g = grid_for((size_t)nnz, 16, kBlock, 2048);
    k_vi_encode<<<g, kBlock, 0, stream()>>>(nnz, a,
                                            reinterpret_cast<const unsigned long long *>(dict),
                                            (int)keys.size(), vi);
  }
  sync();
  return (int)keys.size();
}

// Collects the distinct column offsets (col - row) of a CSR matrix into
// dkeys (ascending, as 32-bit patterns in 64-bit keys).  Returns the count,
// 0 when more than 255 or a row is longer than max_row, -1 after an error.
static int delta_keys(int n, const int *ia, const int *ja, int max_row,
                      std::vector<unsigned long long> &dkeys) {
  static DevBufPtr table, state;
  if (!table) {
    table = dev_alloc(sizeof(unsigned long long) * kViSlots);
    state = dev_alloc(3 * sizeof(int));
    if (!table || !state || !table->p || !state->p) return -1;
  }
  unsigned long long *t = table->as<unsigned long long>();
  int *st = state->as<int>();
  k_vi_fill<<<4, kBlock, 0, stream()>>>(t, st);
  k_delta_collect<<<grid_for((size_t)n, 16, kBlock, 1024), kBlock, 0, stream()>>>(n, ia, ja, t,
                                                                                  st);
  std::vector<unsigned long long> h = dev_download(t, kViSlots);
  std::vector<int> hs = dev_download(st, 3);
  if (lap_last_status() == LAP_ERR_HIP) return -1;
  if (hs[1] || hs[0] > kViMax || hs[2] > max_row) return 0;
  dkeys.clear();
  for (unsigned long long k : h)
    if (k != kViEmpty) dkeys.push_back(k);
  std::sort(dkeys.begin(), dkeys.end());
  return (int)dkeys.size();
}

int dvi_encode(int n, long long nnz, const int *ia, const int *ja, const double *a, int max_row,
               unsigned short *code, double *vdict, int *ddict, int *nv, int *ndl) {
  *nv = *ndl = 0;
  if (n <= 0 || nnz <= 0) return 0;
  std::vector<unsigned long long> dkeys;
  const int nd = delta_keys(n, ia, ja, max_row, dkeys);
  if (nd <= 0) return nd;
  const int nval = vi_encode(nnz, a, nullptr, vdict);  // the value table only
  if (nval <= 0) return nval;
  static DevBufPtr dk;
  if (!dk) dk = dev_alloc(sizeof(unsigned long long) * kViMax);
  if (!dk || !dk->p) return -1;
  std::vector<int> dint(dkeys.size());
  for (size_t i = 0; i < dkeys.size(); ++i) dint[i] = (int)(unsigned)dkeys[i];
  if (!LAP_HIP(hipMemcpyAsync(dk->p, dkeys.data(), dkeys.size() * 8, hipMemcpyHostToDevice,
                              stream())) ||
      !LAP_HIP(hipMemcpyAsync(ddict, dint.data(), dint.size() * sizeof(int),
                              hipMemcpyHostToDevice, stream())))
    return -1;
  k_dvi_encode<<<grid_for((size_t)n, 4, kBlock, 2048), kBlock, 0, stream()>>>(
      n, ia, ja, a, reinterpret_cast<const unsigned long long *>(vdict), nval,
      dk->as<unsigned long long>(), nd, code);
  sync();
  if (lap_last_status() == LAP_ERR_HIP) return -1;
  *nv = nval;
  *ndl = nd;
  return 1;
}

}  // namespace lap
