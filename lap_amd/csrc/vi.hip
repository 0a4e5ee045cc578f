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
    const unsigned long long k = (unsigned long long)__double_as_longlong(a[e]);
    if (k == kViEmpty) {
      over = 1;
      break;
    }
    unsigned h = vi_hash(k, 9);
    for (int probe = 0; probe < kViLds; ++probe, h = (h + 1) & (kViLds - 1)) {
      const unsigned long long cur = set[h];
      if (cur == k) break;
      if (cur == kViEmpty) {
        const unsigned long long prev = atomicCAS(&set[h], kViEmpty, k);
        if (prev == kViEmpty) {
          if (atomicAdd(&count, 1) >= kViMax) over = 1;
          break;
        }
        if (prev == k) break;
      }
    }
    if (over) break;
  }
  __syncthreads();
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

__global__ __launch_bounds__(kBlock) void k_vi_fill(unsigned long long *table, int *state) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < kViSlots; i += gridDim.x * kBlock)
    table[i] = kViEmpty;
  if (blockIdx.x == 0 && threadIdx.x < 2) state[threadIdx.x] = 0;
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
    state = dev_alloc(2 * sizeof(int));
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
  if (nnz > 0) {
    const unsigned g = grid_for((size_t)nnz, 16, kBlock, 2048);
    k_vi_encode<<<g, kBlock, 0, stream()>>>(nnz, a,
                                            reinterpret_cast<const unsigned long long *>(dict),
                                            (int)keys.size(), vi);
  }
  sync();
  return (int)keys.size();
}

}  // namespace lap
