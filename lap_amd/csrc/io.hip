// io.hip -- the input path on the host: Matrix Market reader and writer,
// right-hand-side reader, binary CSR, and definitions of the reference's
// io_utils.h functions (read_mm_file, read_adjacency_file, coo_to_csr,
// read_rhs, split, create_L_and_split).
//
// The reference reads with fgets/sscanf one line at a time
// (io_utils.c:15-130) and sorts every CSR row with qsort (:141-235); at a
// 256^3 grid that is tens of seconds of set-up.  Here the file is mapped,
// cut at line boundaries into one chunk per thread and parsed in parallel
// (integers by hand, values with strtod, the conversion sscanf("%lf")
// performs, so every value keeps its bits), and the chunks are concatenated
// in file order.  COO -> CSR on the host (for the io_utils drop-in) counts
// rows with atomics, scatters each entry with a sort key (column, fill
// sequence) and sorts the rows in parallel; the key makes the result
// independent of thread timing and equal to the reference's, whose fill
// order a stable sort keeps for equal columns.  The device versions of
// COO -> CSR, split and the adjacency Laplacian are in convert.hip.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lap_io.h"
#include "lap_core.hpp"

struct lap_mm {
  lap_mm_info info{};
  std::vector<int> rows, cols;
  std::vector<double> vals;
};

namespace lap {

int host_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(32u, h ? h : 1u));
}

namespace {

// Runs f(i) for i in [0, count) on up to host_threads() threads (i cyclic).
template <class F>
void parallel_for(long long count, F &&f) {
  const int T = (int)std::min<long long>(host_threads(), std::max(count, 1ll));
  if (T == 1) {
    for (long long i = 0; i < count; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t)
    th.emplace_back([&, t]() {
      for (long long i = t; i < count; i += T) f(i);
    });
  for (auto &x : th) x.join();
}

struct Mapped {
  const char *p = nullptr;
  size_t len = 0;
  int fd = -1;
  ~Mapped() {
    if (p) munmap((void *)p, len);
    if (fd >= 0) close(fd);
  }
  bool open_file(const char *path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat st;
    if (fstat(fd, &st) != 0) return false;
    len = (size_t)st.st_size;
    if (len == 0) return true;
    void *m = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) return false;
    p = (const char *)m;
    madvise(m, len, MADV_SEQUENTIAL);
    return true;
  }
};

// The characters isspace() accepts in the C locale, as sscanf skips them.
inline bool blank(char c) { return c == ' ' || (c >= '\t' && c <= '\r' && c != '\n'); }

inline const char *line_end(const char *s, const char *end) {
  const char *e = (const char *)memchr(s, '\n', (size_t)(end - s));
  return e ? e : end;
}

// Signed decimal integer after blanks (sscanf "%d"); nullptr when there is
// none or it does not fit an int.
inline const char *parse_int(const char *s, const char *end, long long &v) {
  while (s < end && blank(*s)) ++s;
  bool neg = false;
  if (s < end && (*s == '-' || *s == '+')) neg = *s++ == '-';
  if (s >= end || *s < '0' || *s > '9') return nullptr;
  long long x = 0;
  while (s < end && *s >= '0' && *s <= '9') {
    x = x * 10 + (*s - '0');
    if (x > 0x7fffffffll + 1) return nullptr;
    ++s;
  }
  v = neg ? -x : x;
  if (v > 0x7fffffffll) return nullptr;
  return s;
}

// Clinger's fast path: a decimal [+-]digits[.digits][(e|E)[+-]digits] whose
// significant digits form an integer m <= 2^53 and whose decimal exponent e
// lies in [-22, 22] is m * 10^e or m / 10^-e with both operands exact in
// double, so one IEEE operation gives the correctly rounded value, which is
// what strtod returns.  Returns nullptr (caller falls back to strtod) for
// anything else, including hexadecimal, inf and nan.
inline const char *parse_decimal_fast(const char *s, const char *end, double &v) {
  static constexpr double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                        1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                        1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
  const char *p = s;
  bool neg = false;
  if (p < end && (*p == '-' || *p == '+')) neg = *p++ == '-';
  uint64_t m = 0;
  int sig = 0, e10 = 0, ndig = 0;
  for (; p < end && *p >= '0' && *p <= '9'; ++p, ++ndig) {
    if (m == 0 && *p == '0') continue;  // leading zeros are not significant
    if (++sig > 19) return nullptr;
    m = m * 10 + (uint64_t)(*p - '0');
  }
  if (p < end && *p == '.') {
    for (++p; p < end && *p >= '0' && *p <= '9'; ++p, ++ndig) {
      --e10;
      if (m == 0 && *p == '0') continue;
      if (++sig > 19) return nullptr;
      m = m * 10 + (uint64_t)(*p - '0');
    }
  }
  if (ndig == 0) return nullptr;
  if (p < end && (*p == 'e' || *p == 'E')) {
    const char *q = p + 1;
    bool eneg = false;
    if (q < end && (*q == '-' || *q == '+')) eneg = *q++ == '-';
    if (q >= end || *q < '0' || *q > '9') return nullptr;
    int x = 0;
    for (; q < end && *q >= '0' && *q <= '9'; ++q)
      if (x < 10000) x = x * 10 + (*q - '0');
    e10 += eneg ? -x : x;
    p = q;
  }
  if (p < end && !blank(*p) && *p != '\n') return nullptr;  // token continues
  if (m > (1ull << 53)) return nullptr;
  double r = (double)m;
  if (m != 0) {
    if (e10 < -22 || e10 > 22) return nullptr;
    r = e10 < 0 ? r / kPow10[-e10] : r * kPow10[e10];
  }
  v = neg ? -r : r;
  return p;
}

// A value after blanks, converted as sscanf "%lf" does (strtod): Clinger's
// fast path when it applies, else strtod on a copy of the token (the mapping
// has no terminator).
inline const char *parse_double(const char *s, const char *end, double &v) {
  while (s < end && blank(*s)) ++s;
  if (const char *q = parse_decimal_fast(s, end, v)) return q;
  char buf[160];
  size_t k = 0;
  while (s + k < end && k < sizeof(buf) - 1 && s[k] != '\n' && !blank(s[k])) {
    buf[k] = s[k];
    ++k;
  }
  if (k == 0) return nullptr;
  buf[k] = 0;
  char *e = nullptr;
  v = strtod(buf, &e);
  if (e == buf) return nullptr;
  return s + (e - buf);
}

// Skips '%' lines and blank lines; returns the start of the next line with
// content (or end).
const char *skip_comments(const char *s, const char *end) {
  while (s < end) {
    const char *le = line_end(s, end);
    const char *t = s;
    while (t < le && blank(*t)) ++t;
    if (t < le && *s != '%') return s;
    s = le + 1;
  }
  return end;
}

struct Chunk {
  std::vector<int> r, c;
  std::vector<double> v;
  long long bad = -1;  // offset of the first malformed line in the chunk
};

void parse_chunk(const char *b, const char *e, bool values, Chunk &out) {
  out.r.reserve((size_t)(e - b) / 16);
  out.c.reserve((size_t)(e - b) / 16);
  if (values) out.v.reserve((size_t)(e - b) / 16);
  for (const char *s = b; s < e;) {
    const char *le = line_end(s, e);
    const char *t = s;
    while (t < le && blank(*t)) ++t;
    if (t < le && *t != '%') {
      long long r = 0, c = 0;
      double v = 1.0;
      const char *q = parse_int(t, le, r);
      if (q) q = parse_int(q, le, c);
      if (q && values) q = parse_double(q, le, v);
      if (!q || r < 1 || c < 1) {
        if (out.bad < 0) out.bad = (long long)(s - b);
      } else {
        out.r.push_back((int)(r - 1));
        out.c.push_back((int)(c - 1));
        if (values) out.v.push_back(v);
      }
    }
    s = le + 1;
  }
}

// Parses path into M (COO in file order).  Returns LAP_OK or an error.
int mm_parse(const char *path, bool adjacency, lap_mm &M) {
  Mapped f;
  if (!path || !f.open_file(path) || !f.p) {
    set_error(LAP_ERR_ARG, "mm_load: cannot read %s (%s)", path ? path : "(null)",
              f.fd < 0 ? strerror(errno) : "empty file");
    return LAP_ERR_ARG;
  }
  const char *p = f.p, *end = f.p + f.len;
  const std::string banner(p, line_end(p, end));
  const bool pattern = banner.find("pattern") != std::string::npos;
  const bool sym = adjacency || banner.find("symmetric") != std::string::npos;
  const char *s = skip_comments(p, end);
  const char *le = line_end(s, end);
  long long n = 0, m = 0, nz = 0;
  const char *q = parse_int(s, le, n);
  if (q) q = parse_int(q, le, m);
  if (q) q = parse_int(q, le, nz);
  if (!q || n < 1 || m < 1 || nz < 0 || (sym && n != m) ||
      nz + (adjacency ? n : 0) > 0x7fffffffll) {
    set_error(LAP_ERR_STRUCT, "mm_load: bad size line in %s", path);
    return LAP_ERR_STRUCT;
  }
  const char *body = le < end ? le + 1 : end;
  const int T = host_threads();
  std::vector<const char *> cut((size_t)T + 1);
  cut[0] = body;
  cut[(size_t)T] = end;
  for (int t = 1; t < T; ++t) {
    const char *c = body + (size_t)(end - body) * (size_t)t / (size_t)T;
    if (c < cut[(size_t)t - 1]) c = cut[(size_t)t - 1];
    const char *e = c < end ? line_end(c, end) : end;
    cut[(size_t)t] = e < end ? e + 1 : end;
  }
  std::vector<Chunk> ch((size_t)T);
  parallel_for(T, [&](long long t) { parse_chunk(cut[t], cut[t + 1], !pattern, ch[t]); });
  long long total = 0;
  for (int t = 0; t < T; ++t) {
    if (ch[(size_t)t].bad >= 0) {
      const char *at = cut[(size_t)t] + ch[(size_t)t].bad;
      set_error(LAP_ERR_STRUCT, "mm_load: malformed entry \"%.*s\" in %s",
                (int)std::min<ptrdiff_t>(line_end(at, end) - at, 60), at, path);
      return LAP_ERR_STRUCT;
    }
    total += (long long)ch[(size_t)t].r.size();
  }
  if (total != nz) {
    set_error(LAP_ERR_STRUCT, "mm_load: %s declares %lld entries and holds %lld", path, nz,
              total);
    return LAP_ERR_STRUCT;
  }
  const size_t count = (size_t)nz + (adjacency ? (size_t)n : 0);
  M.rows.resize(count);
  M.cols.resize(count);
  M.vals.resize(count);
  std::vector<size_t> off((size_t)T + 1, 0);
  for (int t = 0; t < T; ++t) off[(size_t)t + 1] = off[(size_t)t] + ch[(size_t)t].r.size();
  std::atomic<long long> bad_at{-1};
  parallel_for(T, [&](long long t) {
    Chunk &c = ch[(size_t)t];
    const size_t o = off[(size_t)t];
    for (size_t i = 0; i < c.r.size(); ++i)
      if (c.r[i] >= n || c.c[i] >= m) {
        long long want = -1;
        bad_at.compare_exchange_strong(want, (long long)(o + i));
        break;
      }
    std::copy(c.r.begin(), c.r.end(), M.rows.begin() + (ptrdiff_t)o);
    std::copy(c.c.begin(), c.c.end(), M.cols.begin() + (ptrdiff_t)o);
    if (pattern)
      std::fill(M.vals.begin() + (ptrdiff_t)o, M.vals.begin() + (ptrdiff_t)(o + c.r.size()), 1.0);
    else
      std::copy(c.v.begin(), c.v.end(), M.vals.begin() + (ptrdiff_t)o);
    std::vector<int>().swap(c.r);
    std::vector<int>().swap(c.c);
    std::vector<double>().swap(c.v);
  });
  if (bad_at.load() >= 0) {
    const size_t i = (size_t)bad_at.load();
    set_error(LAP_ERR_STRUCT, "mm_load: entry (%d, %d) of %s lies outside %lld x %lld",
              M.rows[i] + 1, M.cols[i] + 1, path, n, m);
    return LAP_ERR_STRUCT;
  }
  if (adjacency)  // diagonal slots of L = D - A (io_utils.c:121-129)
    for (long long j = 0; j < n; ++j) {
      M.rows[(size_t)nz + (size_t)j] = (int)j;
      M.cols[(size_t)nz + (size_t)j] = (int)j;
      M.vals[(size_t)nz + (size_t)j] = 1.0;
    }
  long long mirrored = 0;
  if (sym)
    for (size_t i = 0; i < count; ++i) mirrored += M.rows[i] != M.cols[i];
  if ((long long)count + mirrored > 0x7fffffffll) {
    set_error(LAP_ERR_STRUCT, "mm_load: %lld CSR entries exceed int32",
              (long long)count + mirrored);
    return LAP_ERR_STRUCT;
  }
  M.info.n = (int)n;
  M.info.m = (int)m;
  M.info.nnz = (int)count;
  M.info.nnz_csr = (int)((long long)count + mirrored);
  M.info.symmetric = sym ? 1 : 0;
  M.info.pattern = pattern ? 1 : 0;
  return LAP_OK;
}

// COO -> CSR on the host with the reference's result (see the file
// comment).  ia has n+1 entries, ja/a nnz_csr.
void host_coo_to_csr(int n, long long nnz, const int *rows, const int *cols, const double *vals,
                     bool sym, int *ia, int *ja, double *a) {
  std::vector<int> cnt((size_t)n + 1, 0);
  const long long T = host_threads();
  parallel_for(T, [&](long long t) {
    for (long long i = nnz * t / T; i < nnz * (t + 1) / T; ++i) {
      __atomic_fetch_add(&cnt[(size_t)rows[i]], 1, __ATOMIC_RELAXED);
      if (sym && rows[i] != cols[i]) __atomic_fetch_add(&cnt[(size_t)cols[i]], 1, __ATOMIC_RELAXED);
    }
  });
  ia[0] = 0;
  for (int i = 0; i < n; ++i) ia[i + 1] = ia[i] + cnt[(size_t)i];
  const size_t total = (size_t)ia[n];
  std::vector<int> fill(ia, ia + n);
  std::vector<std::pair<uint64_t, double>> tmp(total);
  parallel_for(T, [&](long long t) {
    for (long long i = nnz * t / T; i < nnz * (t + 1) / T; ++i) {
      const int r = rows[i], c = cols[i];
      const uint32_t seq = (uint32_t)(2 * i);
      int pos = __atomic_fetch_add(&fill[(size_t)r], 1, __ATOMIC_RELAXED);
      tmp[(size_t)pos] = {((uint64_t)(uint32_t)c << 32) | seq, vals[i]};
      if (sym && r != c) {
        pos = __atomic_fetch_add(&fill[(size_t)c], 1, __ATOMIC_RELAXED);
        tmp[(size_t)pos] = {((uint64_t)(uint32_t)r << 32) | (seq + 1), vals[i]};
      }
    }
  });
  const long long nb = T * 16;
  parallel_for(nb, [&](long long b) {
    for (long long r = n * b / nb; r < n * (b + 1) / nb; ++r) {
      std::sort(tmp.begin() + ia[r], tmp.begin() + ia[r + 1],
                [](const std::pair<uint64_t, double> &x, const std::pair<uint64_t, double> &y) {
                  return x.first < y.first;
                });
      for (int k = ia[r]; k < ia[r + 1]; ++k) {
        ja[k] = (int)(tmp[(size_t)k].first >> 32);
        a[k] = tmp[(size_t)k].second;
      }
    }
  });
}

template <class T>
T *calloc_n(size_t count) {
  return (T *)calloc(std::max<size_t>(count, 1), sizeof(T));
}

// split / create_L_and_split on the host: strict L and U in CSR order, D
// with the reference's fields (io_utils.c:268-363).
void host_split(lap_mmatrix *A, lap_mmatrix *L, lap_mmatrix *U, lap_mmatrix *D) {
  const int n = A->n;
  const long long nb = host_threads() * 16ll;
  std::vector<int> cl((size_t)n + 1, 0), cu((size_t)n + 1, 0);
  parallel_for(nb, [&](long long b) {
    for (long long i = n * b / nb; i < n * (b + 1) / nb; ++i) {
      int l = 0, u = 0;
      for (int e = A->csr_ia[i]; e < A->csr_ia[i + 1]; ++e) {
        l += A->csr_ja[e] < i;
        u += A->csr_ja[e] > i;
      }
      cl[(size_t)i + 1] = l;
      cu[(size_t)i + 1] = u;
    }
  });
  for (int i = 0; i < n; ++i) {
    cl[(size_t)i + 1] += cl[(size_t)i];
    cu[(size_t)i + 1] += cu[(size_t)i];
  }
  // the reference sizes L and U as nnz - n and its drivers copy that many
  const size_t want = (size_t)std::max(0, A->nnz - n);
  L->csr_ia = calloc_n<int>((size_t)n + 1);
  U->csr_ia = calloc_n<int>((size_t)n + 1);
  D->csr_ia = calloc_n<int>((size_t)n + 1);
  L->csr_ja = calloc_n<int>(std::max(want, (size_t)cl[(size_t)n]));
  U->csr_ja = calloc_n<int>(std::max(want, (size_t)cu[(size_t)n]));
  D->csr_ja = calloc_n<int>((size_t)n);
  L->csr_vals = calloc_n<double>(std::max(want, (size_t)cl[(size_t)n]));
  U->csr_vals = calloc_n<double>(std::max(want, (size_t)cu[(size_t)n]));
  D->csr_vals = calloc_n<double>((size_t)n);
  parallel_for(nb, [&](long long b) {
    for (long long i = n * b / nb; i < n * (b + 1) / nb; ++i) {
      int pl = cl[(size_t)i], pu = cu[(size_t)i];
      L->csr_ia[i] = pl;
      U->csr_ia[i] = pu;
      for (int e = A->csr_ia[i]; e < A->csr_ia[i + 1]; ++e) {
        const int c = A->csr_ja[e];
        if (c == i) {
          D->csr_vals[i] = A->csr_vals[e];
          D->csr_ia[i] = (int)i;
          D->csr_ja[i] = (int)i;
        } else if (c > i) {
          U->csr_ja[pu] = c;
          U->csr_vals[pu++] = A->csr_vals[e];
        } else {
          L->csr_ja[pl] = c;
          L->csr_vals[pl++] = A->csr_vals[e];
        }
      }
    }
  });
  D->csr_ia[n] = n;
  L->csr_ia[n] = cl[(size_t)n];
  U->csr_ia[n] = cu[(size_t)n];
  L->n = U->n = D->n = n;
  L->m = U->m = D->m = A->m;
  L->nnz = U->nnz = A->nnz - n;
  D->nnz = n;
}

void report_load_failure(const char *fn, lap_mmatrix *A) {
  fprintf(stderr, "%s: %s\n", fn, lap_last_error());
  A->coo_rows = A->coo_cols = nullptr;
  A->coo_vals = nullptr;
  A->n = A->m = A->nnz = 0;
}

void to_mmatrix(lap_mm &M, lap_mmatrix *A) {
  const size_t k = (size_t)M.info.nnz;
  A->coo_rows = calloc_n<int>(k);
  A->coo_cols = calloc_n<int>(k);
  A->coo_vals = calloc_n<double>(k);
  if (k) {
    memcpy(A->coo_rows, M.rows.data(), k * sizeof(int));
    memcpy(A->coo_cols, M.cols.data(), k * sizeof(int));
    memcpy(A->coo_vals, M.vals.data(), k * sizeof(double));
  }
  A->n = M.info.n;
  A->m = M.info.m;
  A->nnz = M.info.nnz;
  A->symmetric = M.info.symmetric;
}

// ------------------------------------------------------------ binary CSR
constexpr char kCsrMagic[8] = {'L', 'A', 'P', 'C', 'S', 'R', 0, 1};
constexpr size_t kCsrHeader = 32;

bool read_exact(int fd, void *dst, size_t bytes, off_t at) {
  char *p = (char *)dst;
  while (bytes) {
    const ssize_t r = pread(fd, p, bytes, at);
    if (r <= 0) return false;
    p += r;
    at += r;
    bytes -= (size_t)r;
  }
  return true;
}

// Reads [at, at+bytes) of fd into host memory with several threads.
bool read_parallel(int fd, char *dst, size_t bytes, off_t at) {
  const size_t piece = 8u << 20;
  const long long pieces = (long long)((bytes + piece - 1) / piece);
  std::atomic<bool> ok{true};
  parallel_for(pieces, [&](long long k) {
    const size_t o = (size_t)k * piece, len = std::min(piece, bytes - o);
    if (!read_exact(fd, dst + o, len, at + (off_t)o)) ok = false;
  });
  return ok.load();
}

// Streams [at, at+bytes) of fd into device memory dst through two pinned
// buffers: while one is copied to the device the other is read.
bool read_to_device(int fd, char *dst, size_t bytes, off_t at) {
  static std::mutex mu;
  static char *pin[2] = {nullptr, nullptr};
  static hipEvent_t ev[2];
  constexpr size_t kPin = 64u << 20;
  std::lock_guard<std::mutex> g(mu);
  if (!pin[0]) {
    for (int k = 0; k < 2; ++k) {
      if (!LAP_HIP(hipHostMalloc((void **)&pin[k], kPin, hipHostMallocDefault))) return false;
      if (!LAP_HIP(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming))) return false;
    }
  }
  bool used[2] = {false, false};
  size_t k = 0;
  for (size_t o = 0; o < bytes; o += kPin, k ^= 1) {
    const size_t len = std::min(kPin, bytes - o);
    if (used[k] && !LAP_HIP(hipEventSynchronize(ev[k]))) return false;
    if (!read_parallel(fd, pin[k], len, at + (off_t)o)) {
      set_error(LAP_ERR_ARG, "csr_load: short read");
      sync();
      return false;
    }
    if (!LAP_HIP(hipMemcpyAsync(dst + o, pin[k], len, hipMemcpyHostToDevice, stream())) ||
        !LAP_HIP(hipEventRecord(ev[k], stream())))
      return false;
    used[k] = true;
  }
  sync();
  return true;
}

}  // namespace

}  // namespace lap

using namespace lap;

extern "C" {

int lap_mm_load(const char *path, int adjacency, lap_mm **out, lap_mm_info *info) {
  if (!path || !out) {
    set_error(LAP_ERR_ARG, "mm_load: null argument");
    return LAP_ERR_ARG;
  }
  *out = nullptr;
  auto M = std::make_unique<lap_mm>();
  const int st = mm_parse(path, adjacency != 0, *M);
  if (st != LAP_OK) return st;
  if (info) *info = M->info;
  *out = M.release();
  return LAP_OK;
}

int lap_mm_coo(const lap_mm *h, int *rows, int *cols, double *vals, int to_device) {
  if (!h || !rows || !cols || !vals) {
    set_error(LAP_ERR_ARG, "mm_coo: null argument");
    return LAP_ERR_ARG;
  }
  const size_t k = (size_t)h->info.nnz;
  if (!to_device) {
    if (k) {
      memcpy(rows, h->rows.data(), k * sizeof(int));
      memcpy(cols, h->cols.data(), k * sizeof(int));
      memcpy(vals, h->vals.data(), k * sizeof(double));
    }
    return LAP_OK;
  }
  if (!check_device("mm_coo", {{rows, "rows", k * sizeof(int)},
                               {cols, "cols", k * sizeof(int)},
                               {vals, "vals", k * sizeof(double)}}))
    return LAP_ERR_ARG;
  if (!k) return LAP_OK;
  bool ok = LAP_HIP(hipMemcpy(rows, h->rows.data(), k * sizeof(int), hipMemcpyHostToDevice));
  ok = ok && LAP_HIP(hipMemcpy(cols, h->cols.data(), k * sizeof(int), hipMemcpyHostToDevice));
  ok = ok && LAP_HIP(hipMemcpy(vals, h->vals.data(), k * sizeof(double), hipMemcpyHostToDevice));
  return ok ? LAP_OK : LAP_ERR_HIP;
}

void lap_mm_free(lap_mm *h) { delete h; }

int lap_rhs_load(const char *path, int n, double *b, int to_device) {
  if (!path || !b || n < 0) {
    set_error(LAP_ERR_ARG, "rhs_load: bad argument");
    return LAP_ERR_ARG;
  }
  Mapped f;
  if (!f.open_file(path) || !f.p) {
    set_error(LAP_ERR_ARG, "rhs_load: cannot read %s", path);
    return LAP_ERR_ARG;
  }
  const char *end = f.p + f.len;
  const char *s = skip_comments(f.p, end);
  const char *le = line_end(s, end);
  long long rows = 0;
  if (!parse_int(s, le, rows) || rows != n) {
    set_error(LAP_ERR_STRUCT, "rhs_load: %s holds %lld rows, %d expected", path, rows, n);
    return LAP_ERR_STRUCT;
  }
  std::vector<double> v((size_t)n);
  s = le < end ? le + 1 : end;
  int i = 0;
  for (; i < n && s < end; s = line_end(s, end) + 1) {
    s = skip_comments(s, end);
    if (s >= end) break;
    if (!parse_double(s, line_end(s, end), v[(size_t)i])) {
      set_error(LAP_ERR_STRUCT, "rhs_load: value %d of %s is not a number", i + 1, path);
      return LAP_ERR_STRUCT;
    }
    ++i;
  }
  if (i != n) {
    set_error(LAP_ERR_STRUCT, "rhs_load: %s holds %d values, %d expected", path, i, n);
    return LAP_ERR_STRUCT;
  }
  if (!to_device) {
    if (n) memcpy(b, v.data(), sizeof(double) * (size_t)n);
    return LAP_OK;
  }
  if (!check_device("rhs_load", {{b, "b", sizeof(double) * (size_t)n}})) return LAP_ERR_ARG;
  return !n || LAP_HIP(hipMemcpy(b, v.data(), sizeof(double) * (size_t)n,
                                 hipMemcpyHostToDevice))
             ? LAP_OK
             : LAP_ERR_HIP;
}

int lap_mm_write(const char *path, int n, const int *ia, const int *ja, const double *a) {
  if (!path || n < 1 || !ia || !ja || !a) {
    set_error(LAP_ERR_ARG, "mm_write: bad argument");
    return LAP_ERR_ARG;
  }
  const long long nb = host_threads() * 8ll;
  std::vector<std::string> part((size_t)nb);
  std::vector<long long> cnt((size_t)nb, 0);
  parallel_for(nb, [&](long long b) {
    std::string &o = part[(size_t)b];
    char line[96];
    for (long long i = n * b / nb; i < n * (b + 1) / nb; ++i)
      for (int e = ia[i]; e < ia[i + 1]; ++e)
        if (ja[e] <= i) {
          const int k = snprintf(line, sizeof line, "%lld %d %.17g\n", i + 1, ja[e] + 1, a[e]);
          o.append(line, (size_t)k);
          ++cnt[(size_t)b];
        }
  });
  long long total = 0;
  for (long long c : cnt) total += c;
  FILE *f = fopen(path, "w");
  if (!f) {
    set_error(LAP_ERR_ARG, "mm_write: cannot write %s (%s)", path, strerror(errno));
    return LAP_ERR_ARG;
  }
  fprintf(f, "%%%%MatrixMarket matrix coordinate real symmetric\n%d %d %lld\n", n, n, total);
  bool ok = true;
  for (const std::string &s : part) ok = ok && fwrite(s.data(), 1, s.size(), f) == s.size();
  ok = (fclose(f) == 0) && ok;
  if (!ok) {
    set_error(LAP_ERR_ARG, "mm_write: short write to %s", path);
    return LAP_ERR_ARG;
  }
  return LAP_OK;
}

int lap_csr_save(const char *path, int n, const int *ia, const int *ja, const double *a) {
  if (!path || n < 0 || !ia || !ja || !a) {
    set_error(LAP_ERR_ARG, "csr_save: bad argument");
    return LAP_ERR_ARG;
  }
  const bool dev = is_device_ptr(ia);
  if (dev != is_device_ptr(ja) || dev != is_device_ptr(a)) {
    set_error(LAP_ERR_ARG, "csr_save: ia, ja and a must all be host or all device memory");
    return LAP_ERR_ARG;
  }
  std::vector<int> hia, hja;
  std::vector<double> ha;
  int nnz;
  if (dev) {
    hia = dev_download(ia, (size_t)n + 1);
    nnz = hia[(size_t)n];
    hja = dev_download(ja, (size_t)nnz);
    ha = dev_download(a, (size_t)nnz);
    ia = hia.data();
    ja = hja.data();
    a = ha.data();
  } else {
    nnz = ia[n];
  }
  if (nnz < 0) {
    set_error(LAP_ERR_STRUCT, "csr_save: ia[n] = %d", nnz);
    return LAP_ERR_STRUCT;
  }
  FILE *f = fopen(path, "wb");
  if (!f) {
    set_error(LAP_ERR_ARG, "csr_save: cannot write %s (%s)", path, strerror(errno));
    return LAP_ERR_ARG;
  }
  const int64_t head[3] = {n, nnz, 0};
  bool ok = fwrite(kCsrMagic, 1, 8, f) == 8 && fwrite(head, 8, 3, f) == 3;
  ok = ok && fwrite(ia, sizeof(int), (size_t)n + 1, f) == (size_t)n + 1;
  ok = ok && fwrite(ja, sizeof(int), (size_t)nnz, f) == (size_t)nnz;
  ok = ok && fwrite(a, sizeof(double), (size_t)nnz, f) == (size_t)nnz;
  ok = (fclose(f) == 0) && ok;
  if (!ok) {
    set_error(LAP_ERR_ARG, "csr_save: short write to %s", path);
    return LAP_ERR_ARG;
  }
  return LAP_OK;
}

int lap_csr_info(const char *path, int *n, int *nnz) {
  if (!path || !n || !nnz) {
    set_error(LAP_ERR_ARG, "csr_info: null argument");
    return LAP_ERR_ARG;
  }
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) {
    set_error(LAP_ERR_ARG, "csr_info: cannot read %s (%s)", path, strerror(errno));
    return LAP_ERR_ARG;
  }
  char head[kCsrHeader];
  struct stat st;
  const bool got = read_exact(fd, head, kCsrHeader, 0) && fstat(fd, &st) == 0;
  close(fd);
  int64_t v[3];
  memcpy(v, head + 8, sizeof v);
  if (!got || memcmp(head, kCsrMagic, 8) != 0 || v[0] < 0 || v[1] < 0 || v[0] > 0x7ffffffe ||
      v[1] > 0x7fffffff ||
      (uint64_t)st.st_size != kCsrHeader + 4 * (uint64_t)(v[0] + 1) + 12 * (uint64_t)v[1]) {
    set_error(LAP_ERR_STRUCT, "csr_info: %s is not a binary CSR file of this library", path);
    return LAP_ERR_STRUCT;
  }
  *n = (int)v[0];
  *nnz = (int)v[1];
  return LAP_OK;
}

int lap_csr_load(const char *path, int *ia, int *ja, double *a, int to_device) {
  int n = 0, nnz = 0;
  const int st = lap_csr_info(path, &n, &nnz);
  if (st != LAP_OK) return st;
  if (!ia || !ja || !a) {
    set_error(LAP_ERR_ARG, "csr_load: null argument");
    return LAP_ERR_ARG;
  }
  if (to_device && !check_device("csr_load", {{ia, "ia", 4 * ((size_t)n + 1)},
                                              {ja, "ja", 4 * (size_t)nnz},
                                              {a, "a", 8 * (size_t)nnz}}))
    return LAP_ERR_ARG;
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) {
    set_error(LAP_ERR_ARG, "csr_load: cannot read %s", path);
    return LAP_ERR_ARG;
  }
  const off_t o_ia = (off_t)kCsrHeader, o_ja = o_ia + 4 * ((off_t)n + 1),
              o_a = o_ja + 4 * (off_t)nnz;
  bool ok;
  if (to_device) {
    ok = read_to_device(fd, (char *)ia, 4 * ((size_t)n + 1), o_ia) &&
         read_to_device(fd, (char *)ja, 4 * (size_t)nnz, o_ja) &&
         read_to_device(fd, (char *)a, 8 * (size_t)nnz, o_a);
    forget_plans(ia);
    forget_plans(ja);
    forget_plans(a);
  } else {
    ok = read_parallel(fd, (char *)ia, 4 * ((size_t)n + 1), o_ia) &&
         read_parallel(fd, (char *)ja, 4 * (size_t)nnz, o_ja) &&
         read_parallel(fd, (char *)a, 8 * (size_t)nnz, o_a);
  }
  close(fd);
  if (!ok) {
    set_error(LAP_ERR_ARG, "csr_load: short read of %s", path);
    return LAP_ERR_ARG;
  }
  return LAP_OK;
}

// ---------------------------------------------------- io_utils.h drop-in

void read_mm_file(const char *matrixFileName, lap_mmatrix *A) {
  lap_mm M;
  if (!A) return;
  if (mm_parse(matrixFileName, false, M) != LAP_OK) return report_load_failure("read_mm_file", A);
  to_mmatrix(M, A);
}

void read_adjacency_file(const char *matrixFileName, lap_mmatrix *A) {
  lap_mm M;
  if (!A) return;
  if (mm_parse(matrixFileName, true, M) != LAP_OK)
    return report_load_failure("read_adjacency_file", A);
  to_mmatrix(M, A);
}

void coo_to_csr(lap_mmatrix *A) {
  long long total = 0;
  for (int i = 0; i < A->nnz; ++i)
    total += 1 + (A->symmetric && A->coo_rows[i] != A->coo_cols[i]);
  A->nnz_unpacked = (int)total;
  A->csr_ia = calloc_n<int>((size_t)A->n + 1);
  A->csr_ja = calloc_n<int>((size_t)total);
  A->csr_vals = calloc_n<double>((size_t)total);
  host_coo_to_csr(A->n, A->nnz, A->coo_rows, A->coo_cols, A->coo_vals, A->symmetric != 0,
                  A->csr_ia, A->csr_ja, A->csr_vals);
  free(A->coo_rows);
  free(A->coo_cols);
  free(A->coo_vals);
  A->coo_rows = A->coo_cols = nullptr;
  A->coo_vals = nullptr;
}

void read_rhs(const char *rhsFileName, double *rhs) {
  // As the reference: as many values as the size line's row count.
  Mapped f;
  if (!rhs || !rhsFileName || !f.open_file(rhsFileName) || !f.p) {
    fprintf(stderr, "read_rhs: cannot read %s\n", rhsFileName ? rhsFileName : "(null)");
    return;
  }
  const char *end = f.p + f.len;
  const char *s = skip_comments(f.p, end);
  const char *le = line_end(s, end);
  long long rows = 0;
  if (!parse_int(s, le, rows)) {
    fprintf(stderr, "read_rhs: bad size line in %s\n", rhsFileName);
    return;
  }
  s = le < end ? le + 1 : end;
  for (long long i = 0; i < rows && s < end; s = line_end(s, end) + 1) {
    s = skip_comments(s, end);
    if (s >= end) break;
    if (parse_double(s, line_end(s, end), rhs[i])) ++i;
  }
}

void split(lap_mmatrix *A, lap_mmatrix *L, lap_mmatrix *U, lap_mmatrix *D) {
  host_split(A, L, U, D);
}

void create_L_and_split(lap_mmatrix *A, lap_mmatrix *L, lap_mmatrix *U, lap_mmatrix *D,
                        int weighted) {
  const int n = A->n;
  std::vector<int> deg((size_t)n);
  const long long nb = host_threads() * 16ll;
  parallel_for(nb, [&](long long b) {
    for (long long i = n * b / nb; i < n * (b + 1) / nb; ++i) {
      int dd = -1;  // the row's own slot is not counted
      for (int e = A->csr_ia[i]; e < A->csr_ia[i + 1]; ++e)
        dd = (int)((double)dd + A->csr_vals[e]);
      deg[(size_t)i] = dd;
    }
  });
  parallel_for(nb, [&](long long b) {
    for (long long i = n * b / nb; i < n * (b + 1) / nb; ++i) {
      const double dsq = weighted ? 1.0 / sqrt((double)deg[(size_t)i]) : 0.0;
      for (int e = A->csr_ia[i]; e < A->csr_ia[i + 1]; ++e) {
        const int c = A->csr_ja[e];
        double &v = A->csr_vals[e];
        if (c == i)
          v = weighted ? 1.0 : (double)deg[(size_t)i];
        else
          v = weighted ? (-1.0 * v) * dsq * (1.0 / sqrt((double)deg[(size_t)c])) : -1.0 * v;
      }
    }
  });
  host_split(A, L, U, D);
}

}  // extern "C"
