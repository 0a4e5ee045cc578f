// convert.hip -- matrix set-up on the device: COO -> CSR (lap_coo_to_csr),
// A = L + D + U with or without the diagonal in both triangles
// (lap_csr_split) and the graph Laplacian of an adjacency matrix
// (lap_adjacency_laplacian).
//
// COO -> CSR follows coo_to_csr (io_utils.c:141-235): rows are counted,
// every entry (and, for a symmetric matrix, its mirror) is scattered into
// its row, and each row is sorted by column.  The scatter uses atomics, so
// the order inside a row is arbitrary; each entry therefore carries the key
// (column << 32 | fill sequence), where the fill sequence is the position
// the reference's serial fill gives it (2*i for COO entry i, 2*i+1 for its
// mirror).  Keys are unique, so the sorted rows do not depend on timing and
// equal a stable sort of the reference's fill, which is the reference's
// result (glibc's qsort is a merge sort, and for distinct columns the order
// is unique anyway).  Rows are sorted by length class:
//   <= 32 entries      one thread per row, insertion sort;
//   <= 4096 entries    one workgroup per row, bitonic sort in LDS;
//   longer             one workgroup per row: 4096-entry runs sorted in LDS,
//                      then merged pairwise (merge path) through global
//                      scratch.
// Everything here is set-up work, run once per matrix; it is HBM- and
// atomic-bound and makes no attempt at a roofline.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/lap_io.h"
#include "lap_core.hpp"

namespace lap {
namespace {

constexpr int kShortRow = 32;
constexpr int kLdsSort = 4096;  // entries per LDS sort (keys 32 KiB + index 8 KiB)

// --------------------------------------------------------------- scans
constexpr int kScanPer = 4;
constexpr int kScanTile = kBlock * kScanPer;

// Exclusive scan of one value per thread over the workgroup; *total gets
// the workgroup's sum.
__device__ long long block_exscan(long long v, long long *lds, long long *total) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  long long x = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const long long y = __shfl_up(x, o, kWave);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) lds[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long s = 0;
    for (int i = 0; i < kBlock / kWave; ++i) {
      const long long t = lds[i];
      lds[i] = s;
      s += t;
    }
    lds[kBlock / kWave] = s;
  }
  __syncthreads();
  const long long r = x - v + lds[w];
  *total = lds[kBlock / kWave];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kBlock) k_tile_sums(const int *in, long long n, long long *sums) {
  __shared__ long long lds[kBlock / kWave + 1];
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanPer;
  long long s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k)
    if (base + k < n) s += in[base + k];
  long long total;
  block_exscan(s, lds, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One workgroup: exclusive scan of sums[0..nb) in place; sums[nb] = total.
__global__ void __launch_bounds__(kBlock) k_scan_sums(long long *sums, int nb) {
  __shared__ long long lds[kBlock / kWave + 1];
  long long carry = 0;
  for (int b0 = 0; b0 < nb; b0 += kBlock) {
    const int i = b0 + (int)threadIdx.x;
    const long long v = i < nb ? sums[i] : 0;
    long long total;
    const long long r = block_exscan(v, lds, &total);
    if (i < nb) sums[i] = carry + r;
    carry += total;
  }
  if (threadIdx.x == 0) sums[nb] = carry;
}

// out[i] = sum of in[0..i) for i <= n (out has n + 1 entries).
__global__ void __launch_bounds__(kBlock) k_scan_tiles(const int *in, long long n,
                                                       const long long *sums, int nb, int *out) {
  __shared__ long long lds[kBlock / kWave + 1];
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanPer;
  int v[kScanPer];
  long long s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    s += v[k];
  }
  long long total;
  long long r = block_exscan(s, lds, &total) + sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    if (base + k < n) out[base + k] = (int)r;
    r += v[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = (int)sums[nb];
}

}  // namespace

// out[0..n] = exclusive scan of in[0..n); returns the total (or -1 after a
// HIP error).  ws needs scan_ws_bytes(n).
size_t scan_ws_bytes(long long n) {
  return sizeof(long long) * (size_t)((n + kScanTile - 1) / kScanTile + 2);
}
long long scan_exclusive(const int *in, long long n, int *out, long long *ws) {
  const int nb = (int)std::max<long long>(1, (n + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(k_tile_sums, dim3(nb), dim3(kBlock), 0, stream(), in, n, ws);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, stream(), ws, nb);
  hipLaunchKernelGGL(k_scan_tiles, dim3(nb), dim3(kBlock), 0, stream(), in, n, ws, nb, out);
  long long total = -1;
  if (!LAP_HIP(hipMemcpyAsync(&total, ws + nb, sizeof total, hipMemcpyDeviceToHost, stream())))
    return -1;
  sync();
  return total;
}

namespace {

__global__ void k_fill_int(long long n, int v, int *dst) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = v;
}

// ---------------------------------------------------------- COO -> CSR
__global__ void k_coo_count(long long nnz, int n, int m, const int *rows, const int *cols,
                            int sym, int *cnt, int *err) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
       i += (long long)gridDim.x * blockDim.x) {
    const int r = rows[i], c = cols[i];
    if (r < 0 || r >= n || c < 0 || c >= m) {
      atomicOr(err, 1);
      continue;
    }
    atomicAdd(&cnt[r], 1);
    if (sym && r != c) atomicAdd(&cnt[c], 1);
  }
}

__global__ void k_coo_scatter(long long nnz, const int *rows, const int *cols, const double *vals,
                              int sym, int *fill, unsigned long long *key, double *tv) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
       i += (long long)gridDim.x * blockDim.x) {
    const int r = rows[i], c = cols[i];
    const double v = vals[i];
    const unsigned seq = (unsigned)(2 * i);
    int pos = atomicAdd(&fill[r], 1);
    key[pos] = ((unsigned long long)(unsigned)c << 32) | seq;
    tv[pos] = v;
    if (sym && r != c) {
      pos = atomicAdd(&fill[c], 1);
      key[pos] = ((unsigned long long)(unsigned)r << 32) | (seq + 1u);
      tv[pos] = v;
    }
  }
}

// Lists rows longer than kShortRow: medium (<= kLdsSort) and long rows,
// long rows with a scratch offset.  cnt: [0] medium, [1] long; off: scratch
// entries handed out so far.
__global__ void k_row_classes(int n, const int *ia, int *med, int *lng, long long *lng_off,
                              int *cnt, unsigned long long *off) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int len = ia[i + 1] - ia[i];
    if (len <= kShortRow) continue;
    if (len <= kLdsSort) {
      med[atomicAdd(&cnt[0], 1)] = i;
    } else {
      const int k = atomicAdd(&cnt[1], 1);
      lng[k] = i;
      lng_off[k] = (long long)atomicAdd(off, (unsigned long long)len);
    }
  }
}

__device__ __forceinline__ int col_of(unsigned long long k) { return (int)(k >> 32); }

__global__ void k_sort_short(int n, const int *ia, unsigned long long *key, double *tv, int *ja,
                             double *a) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int s = ia[i], len = ia[i + 1] - s;
    if (len > kShortRow) continue;
    for (int t = 1; t < len; ++t) {
      const unsigned long long k = key[s + t];
      const double v = tv[s + t];
      int j = t - 1;
      while (j >= 0 && key[s + j] > k) {
        key[s + j + 1] = key[s + j];
        tv[s + j + 1] = tv[s + j];
        --j;
      }
      key[s + j + 1] = k;
      tv[s + j + 1] = v;
    }
    for (int t = 0; t < len; ++t) {
      ja[s + t] = col_of(key[s + t]);
      a[s + t] = tv[s + t];
    }
  }
}

// Sorts k[0..P) ascending with its index array (P a power of two <=
// kLdsSort), all threads of the workgroup.
__device__ void lds_bitonic(unsigned long long *k, unsigned short *ix, int P) {
  for (int w = 2; w <= P; w <<= 1)
    for (int j = w >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += kBlock) {
        const int l = i ^ j;
        if (l > i) {
          const unsigned long long x = k[i], y = k[l];
          const bool up = (i & w) == 0;
          if ((x > y) == up) {
            k[i] = y;
            k[l] = x;
            const unsigned short t = ix[i];
            ix[i] = ix[l];
            ix[l] = t;
          }
        }
      }
      __syncthreads();
    }
}

// Loads key[s..s+len) (len <= kLdsSort) into LDS padded to a power of two
// and sorts it.
__device__ int lds_load_sort(const unsigned long long *key, long long s, int len,
                             unsigned long long *k, unsigned short *ix) {
  int P = 1;
  while (P < len) P <<= 1;
  for (int t = threadIdx.x; t < P; t += kBlock) {
    k[t] = t < len ? key[s + t] : ~0ull;
    ix[t] = (unsigned short)t;
  }
  __syncthreads();
  lds_bitonic(k, ix, P);
  return P;
}

__global__ void __launch_bounds__(kBlock) k_sort_medium(const int *rowlist, const int *ia,
                                                        const unsigned long long *key,
                                                        const double *tv, int *ja, double *a) {
  __shared__ unsigned long long k[kLdsSort];
  __shared__ unsigned short ix[kLdsSort];
  const int row = rowlist[blockIdx.x];
  const int s = ia[row], len = ia[row + 1] - s;
  lds_load_sort(key, s, len, k, ix);
  for (int t = threadIdx.x; t < len; t += kBlock) {
    ja[s + t] = col_of(k[t]);
    a[s + t] = tv[s + ix[t]];
  }
}

// Number of entries of run A among the first d outputs of merge(A, B)
// (keys unique).
__device__ int merge_split(const unsigned long long *A, int la, const unsigned long long *B,
                           int lb, int d) {
  int lo = d > lb ? d - lb : 0, hi = d < la ? d : la;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (A[mid] < B[d - 1 - mid])
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(kBlock) k_sort_long(const int *rowlist, const long long *rowoff,
                                                      const int *ia,
                                                      const unsigned long long *key,
                                                      const double *tv,
                                                      unsigned long long *k0, int *i0,
                                                      unsigned long long *k1, int *i1, int *ja,
                                                      double *a) {
  __shared__ unsigned long long k[kLdsSort];
  __shared__ unsigned short ix[kLdsSort];
  const int row = rowlist[blockIdx.x];
  const long long s = ia[row];
  const int len = ia[row + 1] - (int)s;
  const long long o = rowoff[blockIdx.x];
  unsigned long long *kb[2] = {k0 + o, k1 + o};
  int *ib[2] = {i0 + o, i1 + o};
  // runs of kLdsSort sorted in LDS
  for (int c0 = 0; c0 < len; c0 += kLdsSort) {
    const int cl = min(kLdsSort, len - c0);
    lds_load_sort(key, s + c0, cl, k, ix);
    for (int t = threadIdx.x; t < cl; t += kBlock) {
      kb[0][c0 + t] = k[t];
      ib[0][c0 + t] = c0 + ix[t];
    }
    __syncthreads();
  }
  // pairwise merges; each thread produces a contiguous slice of every merge
  int src = 0;
  for (int w = kLdsSort; w < len; w <<= 1) {
    for (int lo = 0; lo < len; lo += 2 * w) {
      const int la = min(w, len - lo), lb = max(0, min(w, len - lo - w));
      const unsigned long long *A = kb[src] + lo, *B = A + la;
      const int *IA = ib[src] + lo, *IB = IA + la;
      unsigned long long *O = kb[src ^ 1] + lo;
      int *OI = ib[src ^ 1] + lo;
      const int m = la + lb;
      const int d0 = (int)((long long)m * threadIdx.x / kBlock);
      const int d1 = (int)((long long)m * (threadIdx.x + 1) / kBlock);
      int x = merge_split(A, la, B, lb, d0), y = d0 - x;
      for (int d = d0; d < d1; ++d) {
        const bool takeA = y >= lb || (x < la && A[x] < B[y]);
        if (takeA) {
          O[d] = A[x];
          OI[d] = IA[x];
          ++x;
        } else {
          O[d] = B[y];
          OI[d] = IB[y];
          ++y;
        }
      }
    }
    __syncthreads();
    src ^= 1;
  }
  for (int t = threadIdx.x; t < len; t += kBlock) {
    ja[s + t] = col_of(kb[src][t]);
    a[s + t] = tv[s + ib[src][t]];
  }
}

// ---------------------------------------------------------------- split
__global__ void k_split_count(int n, const int *ia, const int *ja, int aug, int *cl, int *cu) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int l = 0, u = 0;
    for (int e = ia[i]; e < ia[i + 1]; ++e) {
      const int c = ja[e];
      l += c < i;
      u += c > i;
    }
    cl[i] = l + aug;
    cu[i] = u + aug;
  }
}

__global__ void k_split_fill(int n, const int *ia, const int *ja, const double *a, int aug,
                             const int *lia, int *lja, double *la, const int *uia, int *uja,
                             double *ua, double *d) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int pl = lia[i], pu = uia[i] + aug;
    double dv = 0.0;
    for (int e = ia[i]; e < ia[i + 1]; ++e) {
      const int c = ja[e];
      if (c < i) {
        lja[pl] = c;
        la[pl++] = a[e];
      } else if (c > i) {
        uja[pu] = c;
        ua[pu++] = a[e];
      } else {
        dv = a[e];
      }
    }
    if (aug) {
      lja[pl] = i;
      la[pl] = dv;
      uja[uia[i]] = i;
      ua[uia[i]] = dv;
    }
    if (d) d[i] = dv;
  }
}

// ------------------------------------------------- adjacency Laplacian
__global__ void k_adj_degree(int n, const int *ia, const double *a, int *deg) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int dd = -1;
    for (int e = ia[i]; e < ia[i + 1]; ++e) dd = (int)((double)dd + a[e]);
    deg[i] = dd;
  }
}

__global__ void k_adj_values(int n, const int *ia, const int *ja, double *a, const int *deg,
                             int weighted) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double dsq = weighted ? 1.0 / sqrt((double)deg[i]) : 0.0;
    for (int e = ia[i]; e < ia[i + 1]; ++e) {
      const int c = ja[e];
      if (c == i)
        a[e] = weighted ? 1.0 : (double)deg[i];
      else
        a[e] = weighted ? (-1.0 * a[e]) * dsq * (1.0 / sqrt((double)deg[c])) : -1.0 * a[e];
    }
  }
}

unsigned rows_grid(long long n) { return grid_for((size_t)std::max<long long>(n, 1), 1, kBlock, 65536); }

}  // namespace
}  // namespace lap

using namespace lap;

extern "C" {

int lap_coo_to_csr(int n, int m, int nnz, const int *rows, const int *cols, const double *vals,
                   int symmetric, int cap, int *ia, int *ja, double *a, int *nnz_csr) {
  if (n < 1 || m < 1 || nnz < 0 || cap < 0 || !nnz_csr || (symmetric && n != m)) {
    set_error(LAP_ERR_ARG, "coo_to_csr: bad sizes (n %d, m %d, nnz %d, cap %d)", n, m, nnz, cap);
    return LAP_ERR_ARG;
  }
  if (!check_device("coo_to_csr", {{rows, "rows", 4 * (size_t)nnz},
                                   {cols, "cols", 4 * (size_t)nnz},
                                   {vals, "vals", 8 * (size_t)nnz},
                                   {ia, "ia", 4 * ((size_t)n + 1)}}))
    return LAP_ERR_ARG;
  DevBuf cnt(4 * ((size_t)n + 1) + 16), ws(scan_ws_bytes(n) + 16);
  int *c = cnt.as<int>();
  hipLaunchKernelGGL(k_fill_int, dim3(rows_grid(n + 1)), dim3(kBlock), 0, stream(),
                     (long long)n + 1, 0, c);
  if (nnz)
    hipLaunchKernelGGL(k_coo_count, dim3(rows_grid(nnz)), dim3(kBlock), 0, stream(),
                       (long long)nnz, n, m, rows, cols, symmetric, c, c + n);
  int err = 0;
  LAP_HIP(hipMemcpyAsync(&err, c + n, sizeof err, hipMemcpyDeviceToHost, stream()));
  sync();
  if (err) {
    set_error(LAP_ERR_STRUCT, "coo_to_csr: an entry lies outside %d x %d", n, m);
    return LAP_ERR_STRUCT;
  }
  const long long total = scan_exclusive(c, n, ia, ws.as<long long>());
  if (total < 0) return LAP_ERR_HIP;
  if (total > 0x7fffffffll) {
    set_error(LAP_ERR_STRUCT, "coo_to_csr: %lld CSR entries exceed int32", total);
    return LAP_ERR_STRUCT;
  }
  *nnz_csr = (int)total;
  if (total > cap) {
    if (cap == 0) return LAP_OK;  // size query
    set_error(LAP_ERR_ARG, "coo_to_csr: %lld CSR entries, capacity %d", total, cap);
    return LAP_ERR_ARG;
  }
  if (total == 0) return LAP_OK;
  if (!check_device("coo_to_csr", {{ja, "ja", 4 * (size_t)total}, {a, "a", 8 * (size_t)total}}))
    return LAP_ERR_ARG;
  // fill pointers, keyed scatter
  LAP_HIP(hipMemcpyAsync(c, ia, 4 * (size_t)n, hipMemcpyDeviceToDevice, stream()));
  DevBuf key(8 * (size_t)total), tv(8 * (size_t)total);
  hipLaunchKernelGGL(k_coo_scatter, dim3(rows_grid(nnz)), dim3(kBlock), 0, stream(),
                     (long long)nnz, rows, cols, vals, symmetric, c, key.as<unsigned long long>(),
                     tv.as<double>());
  // row classes
  DevBuf med(4 * (size_t)n + 16), lng(4 * (size_t)n + 16), lng_off(8 * (size_t)n + 16),
      ctr(32);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kBlock), 0, stream(), 8ll, 0, ctr.as<int>());
  hipLaunchKernelGGL(k_row_classes, dim3(rows_grid(n)), dim3(kBlock), 0, stream(), n, ia,
                     med.as<int>(), lng.as<int>(), lng_off.as<long long>(), ctr.as<int>(),
                     (unsigned long long *)(ctr.as<int>() + 2));
  hipLaunchKernelGGL(k_sort_short, dim3(rows_grid(n)), dim3(kBlock), 0, stream(), n, ia,
                     key.as<unsigned long long>(), tv.as<double>(), ja, a);
  int h[4] = {0, 0, 0, 0};
  LAP_HIP(hipMemcpyAsync(h, ctr.p, sizeof h, hipMemcpyDeviceToHost, stream()));
  sync();
  unsigned long long long_total;
  memcpy(&long_total, h + 2, sizeof long_total);
  if (h[0])
    hipLaunchKernelGGL(k_sort_medium, dim3(h[0]), dim3(kBlock), 0, stream(), med.as<int>(), ia,
                       key.as<unsigned long long>(), tv.as<double>(), ja, a);
  if (h[1]) {
    DevBuf k0(8 * long_total), k1(8 * long_total), i0(4 * long_total), i1(4 * long_total);
    hipLaunchKernelGGL(k_sort_long, dim3(h[1]), dim3(kBlock), 0, stream(), lng.as<int>(),
                       lng_off.as<long long>(), ia, key.as<unsigned long long>(), tv.as<double>(),
                       k0.as<unsigned long long>(), i0.as<int>(), k1.as<unsigned long long>(),
                       i1.as<int>(), ja, a);
    sync();
  }
  sync();
  forget_plans(ia);
  forget_plans(ja);
  forget_plans(a);
  return lap_last_status();
}

int lap_csr_split(int n, const int *ia, const int *ja, const double *a, int augmented, int cap_l,
                  int *lia, int *lja, double *la, int cap_u, int *uia, int *uja, double *ua,
                  double *d, int *nnz_l, int *nnz_u) {
  if (n < 1 || cap_l < 0 || cap_u < 0 || !nnz_l || !nnz_u || (augmented != 0 && augmented != 1)) {
    set_error(LAP_ERR_ARG, "csr_split: bad arguments");
    return LAP_ERR_ARG;
  }
  if (!check_device("csr_split", {{ia, "ia", 4 * ((size_t)n + 1)},
                                  {lia, "lia", 4 * ((size_t)n + 1)},
                                  {uia, "uia", 4 * ((size_t)n + 1)}}))
    return LAP_ERR_ARG;
  const int nnz = dev_download(ia + n, 1)[0];
  if (!check_device("csr_split", {{ja, "ja", 4 * (size_t)nnz}, {a, "a", 8 * (size_t)nnz}}))
    return LAP_ERR_ARG;
  if (d && !check_device("csr_split", {{d, "d", 8 * (size_t)n}})) return LAP_ERR_ARG;
  if (!spmv_plan(n, nnz, ia, ja)) return lap_last_status();  // validates ia / ja
  DevBuf cl(4 * (size_t)n + 16), cu(4 * (size_t)n + 16), ws(scan_ws_bytes(n) + 16);
  hipLaunchKernelGGL(k_split_count, dim3(rows_grid(n)), dim3(kBlock), 0, stream(), n, ia, ja,
                     augmented, cl.as<int>(), cu.as<int>());
  const long long tl = scan_exclusive(cl.as<int>(), n, lia, ws.as<long long>());
  const long long tu = scan_exclusive(cu.as<int>(), n, uia, ws.as<long long>());
  if (tl < 0 || tu < 0) return LAP_ERR_HIP;
  *nnz_l = (int)tl;
  *nnz_u = (int)tu;
  if (cap_l == 0 && cap_u == 0) return LAP_OK;  // size query
  if (tl > cap_l || tu > cap_u) {
    set_error(LAP_ERR_ARG, "csr_split: needs %lld / %lld entries, capacities %d / %d", tl, tu,
              cap_l, cap_u);
    return LAP_ERR_ARG;
  }
  if (!check_device("csr_split", {{lja, "lja", 4 * (size_t)tl},
                                  {la, "la", 8 * (size_t)tl},
                                  {uja, "uja", 4 * (size_t)tu},
                                  {ua, "ua", 8 * (size_t)tu}}))
    return LAP_ERR_ARG;
  hipLaunchKernelGGL(k_split_fill, dim3(rows_grid(n)), dim3(kBlock), 0, stream(), n, ia, ja, a,
                     augmented, lia, lja, la, uia, uja, ua, d);
  sync();
  for (const void *p : {(const void *)lia, (const void *)lja, (const void *)la,
                        (const void *)uia, (const void *)uja, (const void *)ua})
    forget_plans(p);
  return lap_last_status();
}

int lap_adjacency_laplacian(int n, const int *ia, const int *ja, double *a, int weighted,
                            int *deg) {
  if (n < 1 || (weighted != 0 && weighted != 1)) {
    set_error(LAP_ERR_ARG, "adjacency_laplacian: bad arguments");
    return LAP_ERR_ARG;
  }
  if (!check_device("adjacency_laplacian", {{ia, "ia", 4 * ((size_t)n + 1)}})) return LAP_ERR_ARG;
  const int nnz = dev_download(ia + n, 1)[0];
  if (!check_device("adjacency_laplacian",
                    {{ja, "ja", 4 * (size_t)nnz}, {a, "a", 8 * (size_t)nnz}}))
    return LAP_ERR_ARG;
  if (deg && !check_device("adjacency_laplacian", {{deg, "deg", 4 * (size_t)n}}))
    return LAP_ERR_ARG;
  if (!spmv_plan(n, nnz, ia, ja)) return lap_last_status();  // validates ia / ja
  DevBuf own(deg ? 16 : 4 * (size_t)n + 16);
  int *dd = deg ? deg : own.as<int>();
  hipLaunchKernelGGL(k_adj_degree, dim3(rows_grid(n)), dim3(kBlock), 0, stream(), n, ia, a, dd);
  hipLaunchKernelGGL(k_adj_values, dim3(rows_grid(n)), dim3(kBlock), 0, stream(), n, ia, ja, a,
                     dd, weighted);
  sync();
  forget_plans(a);
  return lap_last_status();
}

}  // extern "C"
