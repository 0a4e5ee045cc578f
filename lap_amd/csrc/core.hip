// core.hip -- runtime state, error reporting and the devMem.h surface.
#include <cstdarg>
#include <cstring>

#include "lap_core.hpp"

namespace lap {

namespace {
std::mutex g_mu;
int g_status = LAP_OK;
char g_msg[512] = "";
hipStream_t g_stream = nullptr;
int g_cus = 0;
Options g_opts;
}  // namespace

void set_error(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_status = code;
  snprintf(g_msg, sizeof g_msg, "%s", buf);
  fprintf(stderr, "[lapgfx] error %d: %s\n", code, buf);
}

bool hip_ok(hipError_t e, const char *what, const char *file, int line) {
  if (e == hipSuccess) return true;
  set_error(LAP_ERR_HIP, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  return false;
}

Options &options() { return g_opts; }

hipStream_t stream() {
  if (!g_stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_stream) {
      hipStream_t s;
      if (LAP_HIP(hipStreamCreate(&s))) g_stream = s;  // blocking, default priority
      int dev = 0;
      (void)hipGetDevice(&dev);
      hipDeviceProp_t prop;
      if (hipGetDeviceProperties(&prop, dev) == hipSuccess) g_cus = prop.multiProcessorCount;
    }
  }
  return g_stream;
}

int device_cus() {
  stream();
  return g_cus > 0 ? g_cus : 256;
}

void sync() { LAP_HIP(hipStreamSynchronize(stream())); }

bool is_device_ptr(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
         attr.type == hipMemoryTypeUnified;
}

bool check_device(const char *fn, std::initializer_list<PtrArg> args) {
  for (const PtrArg &a : args) {
    if (!is_device_ptr(a.p)) {
      set_error(LAP_ERR_ARG, "%s: %s (%p) is not device memory", fn, a.name, a.p);
      return false;
    }
    if (a.bytes) {
      hipDeviceptr_t base = nullptr;
      size_t size = 0;
      if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)a.p) == hipSuccess) {
        const size_t avail = (size_t)((const char *)base + size - (const char *)a.p);
        if (avail < a.bytes) {
          set_error(LAP_ERR_ARG, "%s: %s holds %zu bytes, %zu needed", fn, a.name, avail,
                    a.bytes);
          return false;
        }
      } else {
        (void)hipGetLastError();  // extent unknown: accept
      }
    }
  }
  return true;
}

DevBuf::DevBuf(size_t b) : bytes(b) {
  if (b == 0) b = 16;
  if (!LAP_HIP(hipMalloc(&p, b))) p = nullptr;
}
DevBuf::~DevBuf() {
  if (p) (void)hipFree(p);
}
void DevBuf::upload(const void *src, size_t b, size_t off) {
  LAP_HIP(hipMemcpyAsync(static_cast<char *>(p) + off, src, b, hipMemcpyHostToDevice, stream()));
  sync();
}
DevBufPtr dev_alloc(size_t bytes) { return std::make_shared<DevBuf>(bytes); }

namespace {
std::map<std::string, DevBufPtr> g_scratch;
ReduceWs g_ws{};
DevBufPtr g_ws_buf;
}  // namespace

int *pinned_flags() {
  static int *h = nullptr;
  if (!h && !LAP_HIP(hipHostMalloc((void **)&h, 4 * sizeof(int), hipHostMallocDefault)))
    h = nullptr;
  return h;
}

double *scratch_vec(const char *tag, size_t n) {
  auto &slot = g_scratch[tag];
  if (!slot || slot->bytes < n * sizeof(double)) slot = dev_alloc(n * sizeof(double));
  return slot->as<double>();
}

ReduceWs &reduce_ws() {
  if (!g_ws_buf) {
    const size_t pbytes = sizeof(double) * kMaxPartialBlocks * 4;
    const size_t bytes = pbytes + 64 * sizeof(double);
    g_ws_buf = dev_alloc(bytes);
    char *base = g_ws_buf->as<char>();
    LAP_HIP(hipMemsetAsync(base, 0, bytes, stream()));
    sync();
    g_ws.partials = reinterpret_cast<double *>(base);
    g_ws.scalars = reinterpret_cast<double *>(base + pbytes);
  }
  return g_ws;
}

void release_plans();  // defined with the plan registries

}  // namespace lap

using namespace lap;

extern "C" {

int lap_last_status(void) { return g_status; }
const char *lap_last_error(void) { return g_msg; }
void lap_clear_error(void) {
  g_status = LAP_OK;
  g_msg[0] = 0;
}

int lap_set_option(const char *name, int value) {
  Options &o = options();
  int *slot = nullptr;
  if (!strcmp(name, "spmv_rows")) slot = &o.spmv_rows;
  else if (!strcmp(name, "spmv_walk")) slot = &o.spmv_walk;
  else if (!strcmp(name, "vec_grid")) slot = &o.vec_grid;
  else if (!strcmp(name, "graph_iters")) slot = &o.graph_iters;
  else if (!strcmp(name, "use_graph")) slot = &o.use_graph;
  else if (!strcmp(name, "fuse_sweeps")) slot = &o.fuse_sweeps;
  else if (!strcmp(name, "spmv_format")) slot = &o.spmv_format;
  else if (!strcmp(name, "mixed_vectors")) slot = &o.mixed_vectors;
  else if (!strcmp(name, "sptrsv_long")) slot = &o.sptrsv_long;
  else if (!strcmp(name, "sptrsv_mode")) slot = &o.sptrsv_mode;
  else if (!strcmp(name, "spmm_wg_per_cu")) slot = &o.spmm_wg_per_cu;
  else if (!strcmp(name, "spmv_wg_per_cu")) slot = &o.spmv_wg_per_cu;
  if (!slot) return -1;
  if ((!strcmp(name, "spmv_walk") && value != 0 && value != 3) ||
      (!strcmp(name, "vec_grid") && (value < 1 || value > kMaxPartialBlocks / 2))) {
    set_error(LAP_ERR_ARG, "option %s: value %d out of range", name, value);
    return *slot;
  }
  if (!strcmp(name, "spmv_wg_per_cu") && (value < 1 || value > 16)) {
    set_error(LAP_ERR_ARG, "spmv_wg_per_cu must be in [1, 16]");
    return *slot;
  }
  if (!strcmp(name, "spmm_wg_per_cu") && (value < 1 || value > 16)) {
    set_error(LAP_ERR_ARG, "spmm_wg_per_cu must be in [1, 16]");
    return *slot;
  }
  if (!strcmp(name, "sptrsv_mode") && value != 0 && value != 3) {
    set_error(LAP_ERR_ARG, "sptrsv_mode must be 0 or 3");
    return *slot;
  }
  if (!strcmp(name, "sptrsv_long") && (value < 0 || value > 1)) {
    set_error(LAP_ERR_ARG, "sptrsv_long must be 0 or 1");
    return *slot;
  }
  if (!strcmp(name, "mixed_vectors") && (value < 0 || value > 1)) {
    set_error(LAP_ERR_ARG, "mixed_vectors must be 0 or 1");
    return *slot;
  }
  if (!strcmp(name, "spmv_format") && (value < 0 || value > 2)) {
    set_error(LAP_ERR_ARG, "spmv_format must be 0, 1 or 2");
    return *slot;
  }
  if (!strcmp(name, "fuse_sweeps") && (value < 0 || value > 1)) {
    set_error(LAP_ERR_ARG, "fuse_sweeps must be 0 or 1");
    return *slot;
  }
  if (!strcmp(name, "graph_iters") && (value < 1 || value > 4096)) {
    set_error(LAP_ERR_ARG, "graph_iters must be in [1, 4096]");
    return *slot;
  }
  if (!strcmp(name, "use_graph") && (value < 0 || value > 1)) {
    set_error(LAP_ERR_ARG, "use_graph must be 0 or 1");
    return *slot;
  }
  if (!strcmp(name, "spmv_rows") && (value < 1 || value > 2 * kBlock)) {
    set_error(LAP_ERR_ARG, "spmv_rows must be in [1, %d]", 2 * kBlock);
    return *slot;
  }
  const int prev = *slot;
  *slot = value;  // SpMV plans are keyed by the row-block size too
  return prev;
}

void lap_release_all(void) {
  sync();
  release_plans();
  g_scratch.clear();
}

void initialize_handles(void) {
  stream();
  reduce_ws();
  dot_scratch();  // lap_dot_noalloc's scratch on the current device
}

void *mallocForDevice(void *x, int n, int size) {
  (void)x;
  void *p = nullptr;
  const size_t bytes = (size_t)(n > 0 ? n : 0) * (size_t)(size > 0 ? size : 0);
  if (!LAP_HIP(hipMalloc(&p, bytes ? bytes : 16))) return nullptr;
  return p;
}

void memcpyDevice(void *dest, void *src, int n, int size, char *type) {
  const size_t bytes = (size_t)n * (size_t)size;
  if (bytes == 0) return;
  hipMemcpyKind kind;
  bool ok;
  if (!strcmp(type, "H2D")) {
    kind = hipMemcpyHostToDevice;
    ok = check_device("memcpyDevice H2D", {{dest, "dest"}});
  } else if (!strcmp(type, "D2H")) {
    kind = hipMemcpyDeviceToHost;
    ok = check_device("memcpyDevice D2H", {{src, "src"}});
  } else if (!strcmp(type, "D2D")) {
    kind = hipMemcpyDeviceToDevice;
    ok = check_device("memcpyDevice D2D", {{dest, "dest"}, {src, "src"}});
  } else {
    set_error(LAP_ERR_ARG, "memcpyDevice: unknown direction '%s'", type);
    return;
  }
  if (!ok) return;
  LAP_HIP(hipMemcpyAsync(dest, src, bytes, kind, stream()));
  sync();
}

void freeDevice(void *p) {
  sync();
  LAP_HIP(hipFree(p));
}

}  // extern "C"
