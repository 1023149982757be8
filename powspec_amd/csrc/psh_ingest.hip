// psh_ingest.hip -- catalogue ingest at bandwidth: binary SoA files or host
// columns -> device psh_data records, with the per-object weight rule and
// the catalogue sums of the reference's reader (io/read_ascii.c:875-903).
//
// The reference parses ASCII rows into DATA {x[3], w} on the host
// (io/read_ascii.c:533-1142) and hands the array to genr_mesh.  Here the
// objects arrive as columns (x, y, z and optional w_comp, w_fkp, n(z), all
// float32 or all float64), chunk by chunk:
//   host   : read (pread from the file, or copy from the caller's columns)
//            into one of two pinned staging buffers;
//   DMA    : hipMemcpyAsync of the chunk into a device staging buffer on the
//            ingest's own stream;
//   kernel : one thread per object interleaves the columns into psh_data,
//            applies w = w_comp * w_fkp (survey) or w = w_comp (box) and
//            writes per-block FP64 partial sums of w_comp, w^2 and
//            w_comp w_fkp^2 n(z).
// While chunk k is copied and converted, the host reads chunk k+1 into the
// other pinned buffer.  The pipeline runs on a worker thread, so a caller
// can process the previous catalogue on its context's stream meanwhile;
// psh_ingest_wait joins it and adds the partial sums in block order.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "powspec_hip.h"
#include "psh_internal.h"

namespace {

constexpr int kIngThreads = 256;
constexpr int kNCol = 6;   // x, y, z, w_comp, w_fkp, nz
constexpr int kReadThreads = 8;

template <typename T>
struct Cols {
  const T *c[kNCol];   // device staging columns (nullptr: column absent)
};

// Weight rule and sums of io/read_ascii.c:875-903: absent w_comp and w_fkp
// are 1, absent n(z) is 0; boxes keep w = w_comp and sum it only.
template <typename T>
__global__ void __launch_bounds__(kIngThreads)
ingest_kernel(Cols<T> cols, size_t n, int sim, double *__restrict__ out,
              double *__restrict__ partial /*[blocks][3]*/) {
  const size_t i = blockIdx.x * (size_t)kIngThreads + threadIdx.x;
  double swc = 0.0, sw2 = 0.0, sw2n = 0.0;
  if (i < n) {
    const double x = (double)cols.c[0][i], y = (double)cols.c[1][i], z = (double)cols.c[2][i];
    const double wc = cols.c[3] ? (double)cols.c[3][i] : 1.0;
    double w;
    if (sim) {
      w = wc;
    } else {
      const double wf = cols.c[4] ? (double)cols.c[4][i] : 1.0;
      const double nz = cols.c[5] ? (double)cols.c[5][i] : 0.0;
      w = wc * wf;
      sw2 = w * w;
      sw2n = wc * wf * wf * nz;
    }
    swc = wc;
    double2 *o = reinterpret_cast<double2 *>(out + 4 * i);
    o[0] = make_double2(x, y);
    o[1] = make_double2(z, w);
  }
  __shared__ double red[3][kIngThreads / 64];
  for (int o = 32; o > 0; o >>= 1) {
    swc += __shfl_xor(swc, o);
    sw2 += __shfl_xor(sw2, o);
    sw2n += __shfl_xor(sw2n, o);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wid] = swc;
    red[1][wid] = sw2;
    red[2][wid] = sw2n;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0.0;
    for (int w = 0; w < kIngThreads / 64; w++) s += red[threadIdx.x][w];
    partial[3 * (size_t)blockIdx.x + threadIdx.x] = s;
  }
}

}  // namespace

struct psh_ingest {
  int dev = 0;
  size_t chunk = 0;                 // objects per chunk
  hipStream_t stream = nullptr;     // default priority, non-blocking
  void *pinned[2] = {nullptr, nullptr};
  void *stage[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  double *partial = nullptr;        // device, [blocks of all chunks][3]
  size_t partial_cap = 0;           // in blocks
  std::thread worker;
  bool busy = false;
  int rc = PSH_OK;
  std::string err;
  psh_ingest_result res{};
};

namespace {

int ing_err(psh_ingest *g, int rc, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
int ing_err(psh_ingest *g, int rc, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (g) {
    g->rc = rc;
    g->err = buf;
  }
  return rc;
}

#define ICHK(g, call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return ing_err(g, PSH_ERR_DEVICE, "%s: %s", #call, hipGetErrorString(e_));   \
  } while (0)

double now_s() {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

// A source of column chunks: the file (pread) or the caller's host arrays.
struct Source {
  int fd = -1;
  uint64_t n = 0;
  int dtype = 8;
  unsigned cols = 0;
  off_t col_off[kNCol] = {0, 0, 0, 0, 0, 0};   // file offset of each present column
  const void *host[kNCol] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};

// Copy objects [s, s + m) of one column into `d`: pread from the file or
// memcpy from the caller's array.
int read_piece(const Source &src, int c, size_t s, size_t m, char *d) {
  const size_t es = (size_t)src.dtype;
  const size_t bytes = m * es;
  if (src.fd >= 0) {
    size_t got = 0;
    while (got < bytes) {
      const ssize_t r = pread(src.fd, d + got, bytes - got,
                              src.col_off[c] + (off_t)((s * es) + got));
      if (r <= 0) return PSH_ERR_FILE;
      got += (size_t)r;
    }
  } else {
    memcpy(d, static_cast<const char *>(src.host[c]) + s * es, bytes);
  }
  return PSH_OK;
}

// Fill pinned buffer `dst` (column-major, `chunk` objects per column slot)
// with objects [s, s + m), split over host threads (one memory copy per
// thread does not keep up with PCIe).
int read_chunk(psh_ingest *g, const Source &src, size_t s, size_t m, char *dst) {
  const size_t es = (size_t)src.dtype;
  const int ncol = __builtin_popcount(src.cols);
  const size_t total = m * es * (size_t)ncol;
  int nth = (int)(total >> 22);   // one thread per 4 MB
  const int hw = (int)std::thread::hardware_concurrency();
  nth = std::max(1, std::min({nth, kReadThreads, hw > 0 ? hw : 1}));
  std::vector<int> rcs(nth, PSH_OK);
  auto work = [&](int t) {
    for (int c = 0; c < kNCol; c++) {
      if (!(src.cols >> c & 1u)) continue;
      const size_t a = m * (size_t)t / nth, b = m * (size_t)(t + 1) / nth;
      if (b > a && read_piece(src, c, s + a, b - a, dst + (size_t)c * g->chunk * es + a * es))
        rcs[t] = PSH_ERR_FILE;
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nth; t++) th.emplace_back(work, t);
  work(0);
  for (auto &x : th) x.join();
  for (int t = 0; t < nth; t++)
    if (rcs[t]) return ing_err(g, PSH_ERR_FILE, "catalogue file ends early");
  return PSH_OK;
}

int run_pipeline(psh_ingest *g, Source src, int sim, double *d_out) {
  const double t0 = now_s();
  double t_read = 0.0;
  const size_t es = (size_t)src.dtype;
  const size_t nchunks = (src.n + g->chunk - 1) / g->chunk;
  const size_t bpc = (g->chunk + kIngThreads - 1) / kIngThreads;   // blocks per chunk
  if (hipSetDevice(g->dev) != hipSuccess)
    return ing_err(g, PSH_ERR_DEVICE, "cannot select the ingest's device");
  if (nchunks * bpc > g->partial_cap) {
    if (g->partial) (void)hipFree(g->partial);
    g->partial = nullptr;
    g->partial_cap = 0;
    if (hipMalloc((void **)&g->partial, sizeof(double) * 3 * nchunks * bpc) != hipSuccess)
      return ing_err(g, PSH_ERR_MEMORY, "device allocation of the ingest sums failed");
    g->partial_cap = nchunks * bpc;
  }
  for (size_t k = 0; k < nchunks; k++) {
    const int slot = (int)(k & 1);
    const size_t s = k * g->chunk;
    const size_t m = src.n - s < g->chunk ? src.n - s : g->chunk;
    // the pinned buffer of this slot was last read by the DMA of chunk k-2
    ICHK(g, hipEventSynchronize(g->done[slot]));
    const double tr = now_s();
    int rc = read_chunk(g, src, s, m, static_cast<char *>(g->pinned[slot]));
    t_read += now_s() - tr;
    if (rc) return rc;
    Cols<float> cf{};
    Cols<double> cd{};
    for (int c = 0; c < kNCol; c++) {
      if (!(src.cols >> c & 1u)) continue;
      const size_t off = (size_t)c * g->chunk * es;
      ICHK(g, hipMemcpyAsync(static_cast<char *>(g->stage[slot]) + off,
                             static_cast<char *>(g->pinned[slot]) + off, m * es,
                             hipMemcpyHostToDevice, g->stream));
      const char *p = static_cast<const char *>(g->stage[slot]) + off;
      cf.c[c] = reinterpret_cast<const float *>(p);
      cd.c[c] = reinterpret_cast<const double *>(p);
    }
    const unsigned blocks = (unsigned)((m + kIngThreads - 1) / kIngThreads);
    double *part = g->partial + 3 * k * bpc;
    if (src.dtype == 4)
      hipLaunchKernelGGL(ingest_kernel<float>, dim3(blocks), dim3(kIngThreads), 0, g->stream,
                         cf, m, sim, d_out + 4 * s, part);
    else
      hipLaunchKernelGGL(ingest_kernel<double>, dim3(blocks), dim3(kIngThreads), 0, g->stream,
                         cd, m, sim, d_out + 4 * s, part);
    ICHK(g, hipGetLastError());
    ICHK(g, hipEventRecord(g->done[slot], g->stream));
  }
  // per-block partial sums, added in block order on the host (deterministic)
  std::vector<double> h(3 * nchunks * bpc, 0.0);
  for (size_t k = 0; k < nchunks; k++) {
    const size_t m = src.n - k * g->chunk < g->chunk ? src.n - k * g->chunk : g->chunk;
    const size_t blocks = (m + kIngThreads - 1) / kIngThreads;
    ICHK(g, hipMemcpyAsync(h.data() + 3 * k * bpc, g->partial + 3 * k * bpc,
                           sizeof(double) * 3 * blocks, hipMemcpyDeviceToHost, g->stream));
  }
  ICHK(g, hipStreamSynchronize(g->stream));
  psh_ingest_result &r = g->res;
  r.n = src.n;
  r.sumw = r.sumw2 = r.sumw2n = 0.0;
  for (size_t k = 0; k < nchunks; k++) {
    const size_t m = src.n - k * g->chunk < g->chunk ? src.n - k * g->chunk : g->chunk;
    const size_t blocks = (m + kIngThreads - 1) / kIngThreads;
    for (size_t b = 0; b < blocks; b++) {
      const double *p = h.data() + 3 * (k * bpc + b);
      r.sumw += p[0];
      r.sumw2 += p[1];
      r.sumw2n += p[2];
    }
  }
  r.bytes = (double)src.n * es * (double)__builtin_popcount(src.cols);
  r.seconds_read = t_read;
  r.seconds = now_s() - t0;
  return PSH_OK;
}

int start(psh_ingest *g, Source src, int sim, psh_data *d_out, size_t cap) {
  if (g->busy) return ing_err(g, PSH_ERR_ARG, "an ingest is already in flight on this object");
  if (src.n > cap) {
    if (src.fd >= 0) close(src.fd);
    return ing_err(g, PSH_ERR_ARG, "catalogue of %llu objects exceeds the output capacity %zu",
                   (unsigned long long)src.n, cap);
  }
  g->rc = PSH_OK;
  g->err.clear();
  g->res = psh_ingest_result{};
  g->busy = true;
  double *out = reinterpret_cast<double *>(d_out);
  g->worker = std::thread([g, src, sim, out]() {
    const int rc = src.n ? run_pipeline(g, src, sim, out) : PSH_OK;
    if (rc && g->rc == PSH_OK) g->rc = rc;
    if (src.fd >= 0) close(src.fd);
  });
  return PSH_OK;
}

int check_cols(psh_ingest *g, int dtype, unsigned cols) {
  if (dtype != 4 && dtype != 8) return ing_err(g, PSH_ERR_ARG, "column type must be 4 or 8 bytes");
  if ((cols & 7u) != 7u) return ing_err(g, PSH_ERR_ARG, "x, y and z columns are required");
  if (cols >> kNCol) return ing_err(g, PSH_ERR_ARG, "unknown column bits 0x%x", cols);
  return PSH_OK;
}

}  // namespace

extern "C" int psh_ingest_create(size_t chunk_objects, psh_ingest **out) {
  if (!out) return PSH_ERR_ARG;
  *out = nullptr;
  if (chunk_objects == 0) chunk_objects = (size_t)1 << 21;
  psh_ingest *g = new (std::nothrow) psh_ingest();
  if (!g) return PSH_ERR_MEMORY;
  g->chunk = chunk_objects;
  const size_t bytes = chunk_objects * kNCol * sizeof(double);
  if (hipGetDevice(&g->dev) != hipSuccess ||
      hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
    delete g;
    return PSH_ERR_DEVICE;
  }
  for (int s = 0; s < 2; s++) {
    if (hipHostMalloc(&g->pinned[s], bytes, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&g->stage[s], bytes) != hipSuccess ||
        hipEventCreateWithFlags(&g->done[s], hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(g->done[s], g->stream) != hipSuccess) {
      psh_ingest_destroy(g);
      return PSH_ERR_MEMORY;
    }
  }
  *out = g;
  return PSH_OK;
}

extern "C" int psh_ingest_destroy(psh_ingest *g) {
  if (!g) return PSH_OK;
  if (g->worker.joinable()) g->worker.join();
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  for (int s = 0; s < 2; s++) {
    if (g->pinned[s]) (void)hipHostFree(g->pinned[s]);
    if (g->stage[s]) (void)hipFree(g->stage[s]);
    if (g->done[s]) (void)hipEventDestroy(g->done[s]);
  }
  if (g->partial) (void)hipFree(g->partial);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return PSH_OK;
}

extern "C" int psh_cat_write(const char *path, uint64_t n, int dtype, unsigned cols,
                             const void *const col_ptrs[6]) {
  if (!path || !col_ptrs) return PSH_ERR_ARG;
  if (check_cols(nullptr, dtype, cols)) return PSH_ERR_ARG;
  FILE *fp = fopen(path, "wb");
  if (!fp) return PSH_ERR_FILE;
  psh_cat_header h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, PSH_CAT_MAGIC, 8);
  h.n = n;
  h.dtype = (uint32_t)dtype;
  h.cols = cols;
  int rc = fwrite(&h, sizeof h, 1, fp) == 1 ? PSH_OK : PSH_ERR_FILE;
  for (int c = 0; c < kNCol && rc == PSH_OK; c++) {
    if (!(cols >> c & 1u)) continue;
    if (!col_ptrs[c] || fwrite(col_ptrs[c], (size_t)dtype, n, fp) != n) rc = PSH_ERR_FILE;
  }
  if (fclose(fp) != 0) rc = PSH_ERR_FILE;
  return rc;
}

extern "C" int psh_ingest_file(psh_ingest *g, const char *path, int sim, psh_data *d_out,
                               size_t cap) {
  if (!g || !path || !d_out) return ing_err(g, PSH_ERR_ARG, "null argument");
  Source src;
  src.fd = open(path, O_RDONLY);
  if (src.fd < 0) return ing_err(g, PSH_ERR_FILE, "cannot open the catalogue `%s'", path);
  psh_cat_header h;
  struct stat st;
  if (pread(src.fd, &h, sizeof h, 0) != (ssize_t)sizeof h || memcmp(h.magic, PSH_CAT_MAGIC, 8) ||
      fstat(src.fd, &st) != 0) {
    close(src.fd);
    return ing_err(g, PSH_ERR_FILE, "`%s' is not a binary catalogue", path);
  }
  if (check_cols(g, (int)h.dtype, h.cols)) {
    close(src.fd);
    return g->rc;
  }
  src.n = h.n;
  src.dtype = (int)h.dtype;
  src.cols = h.cols;
  off_t off = (off_t)sizeof h;
  for (int c = 0; c < kNCol; c++) {
    if (!(h.cols >> c & 1u)) continue;
    src.col_off[c] = off;
    off += (off_t)(h.n * h.dtype);
  }
  if (off > st.st_size) {
    close(src.fd);
    return ing_err(g, PSH_ERR_FILE, "`%s' is truncated: %lld of %lld bytes", path,
                   (long long)st.st_size, (long long)off);
  }
  return start(g, src, sim, d_out, cap);
}

extern "C" int psh_ingest_host(psh_ingest *g, uint64_t n, int dtype, unsigned cols,
                               const void *const col_ptrs[6], int sim, psh_data *d_out,
                               size_t cap) {
  if (!g || !col_ptrs || !d_out) return ing_err(g, PSH_ERR_ARG, "null argument");
  if (check_cols(g, dtype, cols)) return g->rc;
  Source src;
  src.n = n;
  src.dtype = dtype;
  src.cols = cols;
  for (int c = 0; c < kNCol; c++) {
    if (!(cols >> c & 1u)) continue;
    if (!col_ptrs[c]) return ing_err(g, PSH_ERR_ARG, "column %d is flagged but null", c);
    src.host[c] = col_ptrs[c];
  }
  return start(g, src, sim, d_out, cap);
}

extern "C" int psh_ingest_wait(psh_ingest *g, psh_ingest_result *res) {
  if (!g) return PSH_ERR_ARG;
  if (g->worker.joinable()) g->worker.join();
  g->busy = false;
  if (res) *res = g->res;
  return g->rc;
}

extern "C" const char *psh_ingest_error(const psh_ingest *g) {
  return g ? g->err.c_str() : "null ingest";
}
