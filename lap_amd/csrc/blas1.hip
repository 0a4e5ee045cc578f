// blas1.hip -- dot, axpy, scal and the element-wise ops of hip_blas.h.
//
// All are HBM-streaming kernels: grid-stride loops over n with 64-lane
// coalesced 8-byte accesses, 256-thread blocks, grid capped at 2048 blocks
// (8 per CU) so each lane has several loads in flight.  The dot product is
// per-block partials plus a one-workgroup final sum in a fixed order
// (deterministic for a given n).
#include "lap_core.hpp"

namespace lap {

namespace {

constexpr unsigned kStreamGridCap = 2048;

template <int BS>
__global__ __launch_bounds__(BS) void k_dot_partial(int n, const double *__restrict__ v,
                                                    const double *__restrict__ w,
                                                    double *__restrict__ partials) {
  __shared__ double lds[BS / kWave];
  double s = 0.0;
  const size_t stride = (size_t)gridDim.x * BS;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < (size_t)n; i += stride)
    s += v[i] * w[i];
  const double b = block_sum<BS>(s, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

template <int BS>
__global__ __launch_bounds__(BS) void k_dot_final(int cnt, const double *__restrict__ partials,
                                                  double *__restrict__ out) {
  __shared__ double lds[BS / kWave];
  const double t = sum_partials<BS>(partials, cnt, 1, lds);
  if (threadIdx.x == 0) *out = t;
}

__global__ __launch_bounds__(kBlock) void k_axpy(int n, double alpha,
                                                 const double *__restrict__ x,
                                                 double *__restrict__ y, const int *done) {
  if (done && *done) return;
  const size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < (size_t)n; i += stride)
    y[i] += alpha * x[i];
}

__global__ __launch_bounds__(kBlock) void k_scal(int n, double alpha, double *__restrict__ v,
                                                 const int *done) {
  if (done && *done) return;
  const size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < (size_t)n; i += stride)
    v[i] *= alpha;
}

__global__ __launch_bounds__(kBlock) void k_vec_vec(int n, const double *__restrict__ x,
                                                    const double *__restrict__ y,
                                                    double *__restrict__ r, const int *done) {
  if (done && *done) return;
  const size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < (size_t)n; i += stride)
    r[i] = x[i] * y[i];
}

// in-place safe variant (x or y may alias r)
__global__ __launch_bounds__(kBlock) void k_vec_vec_alias(int n, const double *x,
                                                          const double *y, double *r,
                                                          const int *done) {
  if (done && *done) return;
  const size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < (size_t)n; i += stride)
    r[i] = x[i] * y[i];
}

__global__ __launch_bounds__(kBlock) void k_reciprocal(int n, const double *__restrict__ v,
                                                       double *__restrict__ r) {
  const size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < (size_t)n; i += stride) {
    const double t = v[i];
    r[i] = (t != 0.0) ? 1.0 / t : 0.0;
  }
}

__global__ __launch_bounds__(kBlock) void k_sqrt(int n, const double *__restrict__ v,
                                                 double *__restrict__ r) {
  const size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < (size_t)n; i += stride) {
    const double t = v[i];
    r[i] = (t > 0.0) ? __builtin_sqrt(t) : 0.0;
  }
}

__global__ __launch_bounds__(kBlock) void k_fill(int n, double val, double *__restrict__ r,
                                                 const int *done) {
  if (done && *done) return;
  const size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < (size_t)n; i += stride)
    r[i] = val;
}

__global__ __launch_bounds__(kBlock) void k_copy(int n, const double *__restrict__ s,
                                                 double *__restrict__ d, const int *done) {
  if (done && *done) return;
  const size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < (size_t)n; i += stride)
    d[i] = s[i];
}

inline unsigned sgrid(int n) { return grid_for((size_t)(n > 0 ? n : 1), 4, kBlock, kStreamGridCap); }

}  // namespace

void launch_axpy(int n, double alpha, const double *x, double *y, const int *done) {
  if (n > 0) k_axpy<<<sgrid(n), kBlock, 0, stream()>>>(n, alpha, x, y, done);
}
void launch_scal(int n, double alpha, double *v, const int *done) {
  if (n > 0) k_scal<<<sgrid(n), kBlock, 0, stream()>>>(n, alpha, v, done);
}
void launch_vec_vec(int n, const double *x, const double *y, double *res, const int *done) {
  if (n <= 0) return;
  if (x == res || y == res)
    k_vec_vec_alias<<<sgrid(n), kBlock, 0, stream()>>>(n, x, y, res, done);
  else
    k_vec_vec<<<sgrid(n), kBlock, 0, stream()>>>(n, x, y, res, done);
}
void launch_copy(int n, const double *src, double *dst, const int *done) {
  if (n > 0 && src != dst) k_copy<<<sgrid(n), kBlock, 0, stream()>>>(n, src, dst, done);
}
void launch_fill(int n, double v, double *dst, const int *done) {
  if (n > 0) k_fill<<<sgrid(n), kBlock, 0, stream()>>>(n, v, dst, done);
}
void launch_dot(int n, const double *v, const double *w, double *out) {
  ReduceWs &ws = reduce_ws();
  const unsigned g = sgrid(n);
  k_dot_partial<kBlock><<<g, kBlock, 0, stream()>>>(n > 0 ? n : 0, v, w, ws.partials);
  k_dot_final<kBlock><<<1, kBlock, 0, stream()>>>((int)g, ws.partials, out);
}

// ------------------------------------------------ no-allocation dot product
// lap_dot_noalloc keeps its partial sums in this array instead of the
// library's workspace: one copy per GPU, held until the process ends.  On the
// host `scratch` is a stand-in, not a device address, so host code only ever
// passes the address hipGetSymbolAddress returned for the current device.
constexpr int kNoAllocPartials = (int)kStreamGridCap;  // 2048 doubles, 16 KiB
static __device__ __align__(256) double scratch[kNoAllocPartials];

constexpr int kNoAllocDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kNoAllocDevices];

double *dot_scratch() {
  int dev = 0;
  if (!LAP_HIP(hipGetDevice(&dev))) return nullptr;
  if (dev < 0 || dev >= kNoAllocDevices) {
    set_error(LAP_ERR_ARG, "lap_dot_noalloc: device %d is beyond the %d supported", dev,
              kNoAllocDevices);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error(LAP_ERR_HIP, "lap_dot_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_dev[dev] = static_cast<double *>(ptr);
  }
  return g_scratch_dev[dev];
}

}  // namespace lap

using namespace lap;

extern "C" {

int lap_dot_noalloc(const int n, const double *v, const double *w, double *result) {
  if (n < 0) {
    set_error(LAP_ERR_ARG, "lap_dot_noalloc: n = %d is negative", n);
    return LAP_ERR_ARG;
  }
  const size_t bytes = sizeof(double) * (size_t)n;
  if (!check_device("lap_dot_noalloc", {{result, "result", sizeof(double)}}) ||
      (n > 0 && !check_device("lap_dot_noalloc", {{v, "v", bytes}, {w, "w", bytes}})))
    return LAP_ERR_ARG;
  double *partials = dot_scratch();
  if (!partials) return lap_last_status();
  // hip_dot's kernels and grid, so the result has hip_dot's bits
  const unsigned g = sgrid(n);
  if (g > (unsigned)kNoAllocPartials) {
    set_error(LAP_ERR_ARG, "lap_dot_noalloc: %u partial sums exceed the %d of its scratch", g,
              kNoAllocPartials);
    return LAP_ERR_ARG;
  }
  k_dot_partial<kBlock><<<g, kBlock, 0, stream()>>>(n, v, w, partials);
  k_dot_final<kBlock><<<1, kBlock, 0, stream()>>>((int)g, partials, result);
  if (!LAP_HIP(hipGetLastError()) || !LAP_HIP(hipStreamSynchronize(stream())))
    return LAP_ERR_HIP;
  return LAP_OK;
}

double hip_dot(const int n, const double *v, const double *w) {
  if (n <= 0) return 0.0;
  if (!check_device("hip_dot", {{v, "v"}, {w, "w"}})) return 0.0;
  ReduceWs &ws = reduce_ws();
  launch_dot(n, v, w, ws.scalars + 0);
  double h = 0.0;
  LAP_HIP(hipMemcpyAsync(&h, ws.scalars + 0, sizeof h, hipMemcpyDeviceToHost, stream()));
  sync();
  return h;
}

void hip_scal(const int n, const double alpha, double *v) {
  if (n <= 0 || !check_device("hip_scal", {{v, "v"}})) return;
  launch_scal(n, alpha, v);
  sync();
}

void hip_axpy(const int n, const double alpha, const double *x, double *y) {
  if (n <= 0 || !check_device("hip_axpy", {{x, "x"}, {y, "y"}})) return;
  launch_axpy(n, alpha, x, y);
  sync();
}

void hip_vec_vec(const int n, const double *x, const double *y, double *res) {
  if (n <= 0 || !check_device("hip_vec_vec", {{x, "x"}, {y, "y"}, {res, "res"}})) return;
  launch_vec_vec(n, x, y, res);
  sync();
}

void hip_vector_reciprocal(const int n, const double *v, double *res) {
  if (n <= 0 || !check_device("hip_vector_reciprocal", {{v, "v"}, {res, "res"}})) return;
  k_reciprocal<<<sgrid(n), kBlock, 0, stream()>>>(n, v, res);
  sync();
}

void hip_vector_sqrt(const int n, const double *v, double *res) {
  if (n <= 0 || !check_device("hip_vector_sqrt", {{v, "v"}, {res, "res"}})) return;
  k_sqrt<<<sgrid(n), kBlock, 0, stream()>>>(n, v, res);
  sync();
}

void hip_vec_copy(const int n, const double *src, double *dest) {
  if (n <= 0 || !check_device("hip_vec_copy", {{src, "src"}, {dest, "dest"}})) return;
  LAP_HIP(hipMemcpyAsync(dest, src, sizeof(double) * (size_t)n,
                                    hipMemcpyDeviceToDevice, stream()));
  sync();
}

void hip_vec_zero(const int n, double *vec) {
  if (n <= 0 || !check_device("hip_vec_zero", {{vec, "vec"}})) return;
  launch_fill(n, 0.0, vec);
  sync();
}

void hip_vec_set(const int n, double value, double *vec) {
  if (n <= 0 || !check_device("hip_vec_set", {{vec, "vec"}})) return;
  launch_fill(n, value, vec);
  sync();
}

// blas.h forwarding (blas.c:15-186 with -DHIP=1)
double dot(const int n, const double *v, const double *w) { return hip_dot(n, v, w); }
void axpy(const int n, const double alpha, double *x, double *y) { hip_axpy(n, alpha, x, y); }
void scal(const int n, const double alpha, double *v) { hip_scal(n, alpha, v); }
void vec_vec(const int n, const double *x, double *y, double *res) { hip_vec_vec(n, x, y, res); }
void vector_reciprocal(const int n, const double *v, double *res) {
  hip_vector_reciprocal(n, v, res);
}
void vector_sqrt(const int n, const double *v, double *res) { hip_vector_sqrt(n, v, res); }
void vec_copy(const int n, double *src, double *dest) { hip_vec_copy(n, src, dest); }
void vec_zero(const int n, double *vec) { hip_vec_zero(n, vec); }

}  // extern "C"
