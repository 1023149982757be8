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
#include <cctype>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>
#include <vector>

#include "powspec_hip.h"
#include "psh_decimal.h"
#include "psh_internal.h"
#include "psh_expr.h"

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
__device__ __forceinline__ void emit(double x, double y, double z, double wc, double wf,
                                     double nz, int sim, double *o4, double &swc,
                                     double &sw2, double &sw2n) {
  double w;
  if (sim) {
    w = wc;
  } else {
    w = wc * wf;
    sw2 = w * w;
    sw2n = wc * wf * wf * nz;
  }
  swc = wc;
  double2 *o = reinterpret_cast<double2 *>(o4);
  o[0] = make_double2(x, y);
  o[1] = make_double2(z, w);
}

// Block sums of the three catalogue sums into partial[blockIdx.x][3].
__device__ __forceinline__ void block_sums(double swc, double sw2, double sw2n,
                                           double *__restrict__ partial) {
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

template <typename T>
__global__ void __launch_bounds__(kIngThreads)
ingest_kernel(Cols<T> cols, size_t n, int sim, double *__restrict__ out,
              double *__restrict__ partial /*[blocks][3]*/) {
  const size_t i = blockIdx.x * (size_t)kIngThreads + threadIdx.x;
  double swc = 0.0, sw2 = 0.0, sw2n = 0.0;
  if (i < n)
    emit((double)cols.c[0][i], (double)cols.c[1][i], (double)cols.c[2][i],
         cols.c[3] ? (double)cols.c[3][i] : 1.0, cols.c[4] ? (double)cols.c[4][i] : 1.0,
         cols.c[5] ? (double)cols.c[5][i] : 0.0, sim, out + 4 * i, swc, sw2, sw2n);
  block_sums(swc, sw2, sw2n, partial);
}

// ---- ASCII: one chunk of whole lines ---------------------------------------
constexpr int kSeg = 64;          // bytes per thread when looking for newlines
constexpr int kMaxFields = 32;    // formatter conversions (after truncation)

// Field kinds of the formatter (io/ascii_fmtr.c conversions, restricted).
enum : signed char {
  F_SKIP_ANY = -3,    // %*s-like: any token
  F_SKIP_NUM = -2,    // %*lf / %*d: a number, discarded
  F_UNUSED = -1,      // %lf whose column no expression uses
};

struct AsciiArgs {
  int nf;                       // fields parsed per line
  int comment;                  // comment character, 0: none
  int has;                      // output columns present (bit c)
  // output column 0..5 (expressions: the value slot 0..15), or an F_* kind
  signed char map[kMaxFields];
  unsigned char integer[kMaxFields];   // %d-type: the token must be an integer
  // expressions: program c of the instruction buffer (outputs 0..5, the
  // selection 6) is [pstart[c], pstart[c] + plen[c]); plen 0: absent
  int expr;
  short pstart[7], plen[7];
};

__device__ __forceinline__ bool is_space(char c) {
  return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f' || c == '\n';
}

__global__ void __launch_bounds__(kIngThreads)
count_nl_kernel(const char *__restrict__ buf, size_t nb, uint32_t *__restrict__ cnt,
                size_t nseg) {
  const size_t t = blockIdx.x * (size_t)kIngThreads + threadIdx.x;
  if (t >= nseg) return;
  const size_t a = t * kSeg, b = min(nb, a + kSeg);
  uint32_t c = 0;
  for (size_t i = a; i < b; i++) c += buf[i] == '\n';
  cnt[t] = c;
}

// starts[o + 1] = one past the o-th newline; starts[0] = 0.
__global__ void __launch_bounds__(kIngThreads)
line_start_kernel(const char *__restrict__ buf, size_t nb, const uint32_t *__restrict__ off,
                  uint32_t *__restrict__ starts, size_t nseg) {
  const size_t t = blockIdx.x * (size_t)kIngThreads + threadIdx.x;
  if (t == 0) starts[0] = 0;
  if (t >= nseg) return;
  const size_t a = t * kSeg, b = min(nb, a + kSeg);
  uint32_t o = off[t];
  for (size_t i = a; i < b; i++)
    if (buf[i] == '\n') starts[++o] = (uint32_t)(i + 1);
}

// One thread per line: leading whitespace, comment and empty lines
// (io/read_ascii.c:773-777), then the formatter's fields as whitespace-
// separated tokens; a token that is not a number where one is required
// marks the line bad (the reference's sscanf failure, :799-805).
__global__ void __launch_bounds__(kIngThreads)
parse_kernel(const char *__restrict__ buf, const uint32_t *__restrict__ starts,
             uint32_t nlines, AsciiArgs a, const pshexpr::Ins *__restrict__ prog,
             double *__restrict__ cols, uint32_t *__restrict__ valid, int *__restrict__ err) {
  const uint32_t o = blockIdx.x * kIngThreads + threadIdx.x;
  if (o >= nlines) return;
  const char *p = buf + starts[o];
  const char *e = buf + starts[o + 1] - 1;   // the newline
  while (p < e && is_space(*p)) p++;
  if (p == e || (a.comment && *p == (char)a.comment)) {
    valid[o] = 0;
    return;
  }
  double v[6] = {0, 0, 0, 0, 0, 0};
  double slot[pshexpr::kMaxSlots];   // expressions: the columns they name
  bool ok = true;
  for (int f = 0; f < a.nf && ok; f++) {
    while (p < e && is_space(*p)) p++;
    if (p == e) {
      ok = false;
      break;
    }
    const char *t = p;
    while (p < e && !is_space(*p)) p++;
    const int m = a.map[f];
    if (m == F_SKIP_ANY) continue;
    if (a.integer[f]) {
      const char *q = t;
      if (q < p && (*q == '+' || *q == '-')) q++;
      if (q == p) ok = false;
      for (; q < p && ok; q++) ok = pshdec::is_digit(*q);
      if (!ok) break;
    }
    double x;
    if (!pshdec::parse_double(t, p, &x)) {
      ok = false;
      break;
    }
    if (m >= 0) {
      if (a.expr) slot[m] = x;
      else v[m] = x;
    }
  }
  if (!ok) {
    atomicMin(err, (int)o);
    valid[o] = 0;
    return;
  }
  if (a.expr) {
    // the selection first (io/read_ascii.c:857-866): a line it rejects is
    // dropped, not an error
    if (a.plen[6] && pshexpr::eval(prog + a.pstart[6], a.plen[6], slot) == 0.0) {
      valid[o] = 0;
      return;
    }
    for (int c = 0; c < 6; c++)
      if (a.plen[c]) v[c] = pshexpr::eval(prog + a.pstart[c], a.plen[c], slot);
  }
  valid[o] = 1;
#pragma unroll
  for (int c = 0; c < 6; c++)
    if (a.has >> c & 1) cols[(size_t)c * nlines + o] = v[c];
}

__global__ void __launch_bounds__(kIngThreads)
ascii_emit_kernel(const double *__restrict__ cols, uint32_t nlines, int has,
                  const uint32_t *__restrict__ valid, const uint32_t *__restrict__ idx, int sim,
                  double *__restrict__ out, double *__restrict__ partial) {
  const uint32_t o = blockIdx.x * kIngThreads + threadIdx.x;
  double swc = 0.0, sw2 = 0.0, sw2n = 0.0;
  if (o < nlines && valid[o]) {
    const size_t n = nlines;
    emit(cols[o], cols[n + o], cols[2 * n + o], (has >> 3 & 1) ? cols[3 * n + o] : 1.0,
         (has >> 4 & 1) ? cols[4 * n + o] : 1.0, (has >> 5 & 1) ? cols[5 * n + o] : 0.0, sim,
         out + 4 * (size_t)idx[o], swc, sw2, sw2n);
  }
  block_sums(swc, sw2, sw2n, partial);
}

__global__ void set_int_kernel(int *p, int v) { *p = v; }

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
  // ASCII chunk buffers (device), grown on demand
  uint32_t *seg = nullptr, *segoff = nullptr, *starts = nullptr, *valid = nullptr,
           *idx = nullptr, *scan_tmp = nullptr;
  size_t seg_cap = 0, segoff_cap = 0, starts_cap = 0, valid_cap = 0, idx_cap = 0,
         scan_cap = 0, acols_cap = 0, aerr_cap = 0;
  double *acols = nullptr;
  int *aerr = nullptr;
  pshexpr::Ins *dprog = nullptr;    // expression programs of the ASCII ingest
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

// Room for `blocks` x 3 partial sums.
int ensure_partial(psh_ingest *g, size_t blocks) {
  if (blocks <= g->partial_cap) return PSH_OK;
  if (g->partial) (void)hipFree(g->partial);
  g->partial = nullptr;
  g->partial_cap = 0;
  if (hipMalloc((void **)&g->partial, sizeof(double) * 3 * blocks) != hipSuccess)
    return ing_err(g, PSH_ERR_MEMORY, "device allocation of the ingest sums failed");
  g->partial_cap = blocks;
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
  int prc = ensure_partial(g, nchunks * bpc);
  if (prc) return prc;
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

// ---- ASCII ----------------------------------------------------------------

template <class T>
int grow(psh_ingest *g, T **p, size_t &cap, size_t need) {
  if (need <= cap) return PSH_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  cap = 0;
  const size_t n = need + need / 4 + 1024;
  if (hipMalloc((void **)p, sizeof(T) * n) != hipSuccess)
    return ing_err(g, PSH_ERR_MEMORY, "device allocation for the ASCII parser failed");
  cap = n;
  return PSH_OK;
}

struct AsciiJob {
  int fd = -1;
  long skip = 0;
  AsciiArgs args{};
  std::string path;
  std::vector<pshexpr::Ins> prog;   // expression programs (empty: plain columns)
};

// Reads [off, off + len) of the file into dst with the read threads.
ssize_t pread_par(int fd, char *dst, size_t len, off_t off) {
  const int hw = (int)std::thread::hardware_concurrency();
  int nth = std::max(1, std::min({(int)(len >> 22), kReadThreads, hw > 0 ? hw : 1}));
  std::vector<ssize_t> got(nth, 0);
  auto work = [&](int t) {
    const size_t a = len * (size_t)t / nth, b = len * (size_t)(t + 1) / nth;
    size_t k = a;
    while (k < b) {
      const ssize_t r = pread(fd, dst + k, b - k, off + (off_t)k);
      if (r <= 0) break;
      k += (size_t)r;
    }
    got[t] = (ssize_t)(k - a);
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nth; t++) th.emplace_back(work, t);
  work(0);
  for (auto &x : th) x.join();
  // bytes available from the start without a gap
  ssize_t tot = 0;
  for (int t = 0; t < nth; t++) {
    const size_t a = len * (size_t)t / nth, b = len * (size_t)(t + 1) / nth;
    tot += got[t];
    if ((size_t)got[t] < b - a) break;
  }
  return tot;
}

int run_ascii(psh_ingest *g, AsciiJob job, int sim, double *d_out, size_t cap) {
  const double t0 = now_s();
  double t_read = 0.0;
  if (hipSetDevice(g->dev) != hipSuccess)
    return ing_err(g, PSH_ERR_DEVICE, "cannot select the ingest's device");
  const size_t PB = g->chunk * kNCol * sizeof(double);   // pinned bytes per slot
  std::vector<char> carry;
  off_t foff = 0;
  long to_skip = job.skip;
  uint64_t line_base = 0;   // file lines before the current chunk
  size_t out_n = 0;
  bool eof = false;
  psh_ingest_result &r = g->res;
  r = psh_ingest_result{};
  std::vector<double> hp;
  int rc;
  if (!job.prog.empty()) {
    if (!g->dprog && hipMalloc((void **)&g->dprog, sizeof(pshexpr::Ins) * 7 * pshexpr::kMaxIns))
      return ing_err(g, PSH_ERR_MEMORY, "cannot allocate the expression programs");
    ICHK(g, hipMemcpyAsync(g->dprog, job.prog.data(), sizeof(pshexpr::Ins) * job.prog.size(),
                           hipMemcpyHostToDevice, g->stream));
  }
  for (int k = 0; !eof; k++) {
    const int slot = k & 1;
    ICHK(g, hipEventSynchronize(g->done[slot]));
    char *buf = static_cast<char *>(g->pinned[slot]);
    if (carry.size() >= PB - 1)
      return ing_err(g, PSH_ERR_FILE, "a line of `%s' is longer than the chunk (%zu bytes)",
                     job.path.c_str(), PB);
    memcpy(buf, carry.data(), carry.size());
    const size_t want = PB - 1 - carry.size();
    const double tr = now_s();
    const ssize_t got = pread_par(job.fd, buf + carry.size(), want, foff);
    t_read += now_s() - tr;
    if (got < 0) return ing_err(g, PSH_ERR_FILE, "cannot read `%s'", job.path.c_str());
    foff += got;
    size_t total = carry.size() + (size_t)got;
    r.bytes += (double)got;
    if ((size_t)got < want) {   // end of file: the last line gets its newline
      eof = true;
      if (total && buf[total - 1] != '\n') buf[total++] = '\n';
    }
    // whole lines only; the tail waits for the next chunk
    size_t cut = total;
    while (cut > 0 && buf[cut - 1] != '\n') cut--;
    if (cut == 0 && !eof)
      return ing_err(g, PSH_ERR_FILE, "a line of `%s' is longer than the chunk (%zu bytes)",
                     job.path.c_str(), PB);
    carry.assign(buf + cut, buf + total);
    // header lines (io/read_ascii.c:766-769), counted on the host
    size_t pos = 0;
    while (to_skip > 0 && pos < cut) {
      const char *nl = static_cast<const char *>(memchr(buf + pos, '\n', cut - pos));
      pos = (size_t)(nl - buf) + 1;
      to_skip--;
      line_base++;
    }
    const size_t nb = cut - pos;
    if (nb == 0) {
      ICHK(g, hipEventRecord(g->done[slot], g->stream));
      continue;
    }
    char *dbuf = static_cast<char *>(g->stage[slot]);
    ICHK(g, hipMemcpyAsync(dbuf, buf + pos, nb, hipMemcpyHostToDevice, g->stream));
    ICHK(g, hipEventRecord(g->done[slot], g->stream));   // pinned slot free after this
    const size_t nseg = (nb + kSeg - 1) / kSeg;
    if ((rc = grow(g, &g->seg, g->seg_cap, nseg + 1)) ||
        (rc = grow(g, &g->segoff, g->segoff_cap, nseg + 1)) ||
        (rc = grow(g, &g->scan_tmp, g->scan_cap, psh_assign_scan_bytes(nseg) / 4 + 64)))
      return rc;
    const unsigned sb = (unsigned)((nseg + kIngThreads - 1) / kIngThreads);
    hipLaunchKernelGGL(count_nl_kernel, dim3(sb), dim3(kIngThreads), 0, g->stream, dbuf, nb,
                       g->seg, nseg);
    ICHK(g, psh_exclusive_scan_u32(g->seg, g->segoff, g->scan_tmp, nseg, g->stream));
    uint32_t nlines = 0;
    ICHK(g, hipMemcpyAsync(&nlines, g->segoff + nseg, sizeof nlines, hipMemcpyDeviceToHost,
                           g->stream));
    ICHK(g, hipStreamSynchronize(g->stream));
    if ((rc = grow(g, &g->starts, g->starts_cap, (size_t)nlines + 1)) ||
        (rc = grow(g, &g->valid, g->valid_cap, (size_t)nlines + 1)) ||
        (rc = grow(g, &g->idx, g->idx_cap, (size_t)nlines + 1)) ||
        (rc = grow(g, &g->acols, g->acols_cap, (size_t)6 * nlines + 1)) ||
        (rc = grow(g, &g->aerr, g->aerr_cap, 1)) ||
        (rc = grow(g, &g->scan_tmp, g->scan_cap, psh_assign_scan_bytes(nlines) / 4 + 64)))
      return rc;
    hipLaunchKernelGGL(line_start_kernel, dim3(sb), dim3(kIngThreads), 0, g->stream, dbuf, nb,
                       g->segoff, g->starts, nseg);
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, g->stream, g->aerr, 0x7fffffff);
    const unsigned lb = (unsigned)((nlines + kIngThreads - 1) / kIngThreads);
    hipLaunchKernelGGL(parse_kernel, dim3(lb), dim3(kIngThreads), 0, g->stream, dbuf, g->starts,
                       nlines, job.args, g->dprog, g->acols, g->valid, g->aerr);
    ICHK(g, psh_exclusive_scan_u32(g->valid, g->idx, g->scan_tmp, nlines, g->stream));
    uint32_t nvalid = 0;
    int bad = 0;
    ICHK(g, hipMemcpyAsync(&nvalid, g->idx + nlines, sizeof nvalid, hipMemcpyDeviceToHost,
                           g->stream));
    ICHK(g, hipMemcpyAsync(&bad, g->aerr, sizeof bad, hipMemcpyDeviceToHost, g->stream));
    ICHK(g, hipStreamSynchronize(g->stream));
    if (bad != 0x7fffffff)
      return ing_err(g, PSH_ERR_FILE, "line %llu of `%s' does not match the formatter",
                     (unsigned long long)(line_base + (uint64_t)bad + 1), job.path.c_str());
    if (out_n + nvalid > cap)
      return ing_err(g, PSH_ERR_ARG, "catalogue `%s' holds more than the output capacity %zu",
                     job.path.c_str(), cap);
    const size_t bpc = lb;
    if ((rc = ensure_partial(g, bpc))) return rc;
    hipLaunchKernelGGL(ascii_emit_kernel, dim3(lb), dim3(kIngThreads), 0, g->stream, g->acols,
                       nlines, job.args.has, g->valid, g->idx, sim, d_out + 4 * out_n,
                       g->partial);
    ICHK(g, hipGetLastError());
    hp.resize(3 * bpc);
    ICHK(g, hipMemcpyAsync(hp.data(), g->partial, sizeof(double) * 3 * bpc,
                           hipMemcpyDeviceToHost, g->stream));
    ICHK(g, hipStreamSynchronize(g->stream));
    for (size_t b = 0; b < bpc; b++) {
      r.sumw += hp[3 * b];
      r.sumw2 += hp[3 * b + 1];
      r.sumw2n += hp[3 * b + 2];
    }
    out_n += nvalid;
    line_base += nlines;
  }
  r.n = out_n;
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
  if (g->dprog) (void)hipFree(g->dprog);
  for (void *p : {(void *)g->seg, (void *)g->segoff, (void *)g->starts, (void *)g->valid,
                  (void *)g->idx, (void *)g->scan_tmp, (void *)g->acols, (void *)g->aerr})
    if (p) (void)hipFree(p);
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

// The expressions of DATA_POSITION, DATA_WT_COMP, DATA_WT_FKP, DATA_NZ and
// DATA_SELECTION (io/read_ascii.c:656-735 builds them as libast trees) as
// postfix programs over the formatter's value columns; the fields after
// the last column they name are not read.
static int compile_exprs(psh_ingest *g, const std::vector<int> &kind,
                         const std::vector<int> &vfield, const psh_ascii_expr *ex,
                         AsciiArgs &a, std::vector<pshexpr::Ins> &prog) {
  std::vector<char> integer(vfield.size());
  for (size_t i = 0; i < vfield.size(); i++) integer[i] = kind[vfield[i]] == 1;
  std::vector<int> slots;   // value column of each slot
  const char *src[7] = {ex->pos[0], ex->pos[1], ex->pos[2], ex->wcomp, ex->wfkp, ex->nz,
                        ex->sel};
  static const char *name[7] = {"position 1", "position 2", "position 3", "completeness weight",
                                "FKP weight", "number density", "selection"};
  a.has = 0;
  prog.clear();
  for (int c = 0; c < 7; c++) {
    a.pstart[c] = (short)prog.size();
    a.plen[c] = 0;
    if (!src[c]) {
      if (c < 3) return ing_err(g, PSH_ERR_ARG, "the %s expression is not given", name[c]);
      continue;
    }
    pshexpr::Compiler cp;
    cp.boolean = c == 6;
    cp.integer = &integer;
    cp.slots = &slots;
    if (!cp.compile(src[c]))
      return ing_err(g, PSH_ERR_ARG, "%s expression: %s", name[c], cp.err.c_str());
    prog.insert(prog.end(), cp.out.begin(), cp.out.end());
    a.plen[c] = (short)cp.out.size();
    if (c < 6) a.has |= 1 << c;
  }
  int last = -1;
  for (int col : slots) last = std::max(last, vfield[col]);
  a.nf = last + 1;
  if (a.nf > kMaxFields) return ing_err(g, PSH_ERR_ARG, "more than %d formatter fields", kMaxFields);
  for (int i = 0; i < a.nf; i++) {
    a.map[i] = kind[i] == 3 ? F_SKIP_ANY : (kind[i] == 2 ? F_SKIP_NUM : F_UNUSED);
    a.integer[i] = kind[i] == 1;
  }
  for (size_t sl = 0; sl < slots.size(); sl++) a.map[vfield[slots[sl]]] = (signed char)sl;
  a.expr = 1;
  return PSH_OK;
}

// Formatter of io/ascii_fmtr.c, restricted to whitespace-separated
// conversions: %lf %le %lg %lE %lG %lF (double columns), %d %i %ld %li %u
// %lu (integer columns, read as doubles), %*<conversion> (skipped; a numeric
// conversion must still parse) and %*s / %*c / %*[..] (any token).
static int parse_fmtr(psh_ingest *g, const psh_ascii_fmt *f, AsciiArgs &a,
                      const psh_ascii_expr *ex = nullptr,
                      std::vector<pshexpr::Ins> *prog = nullptr) {
  const char *p = f->fmtr;
  std::vector<int> kind;   // per field: 0 double, 1 integer, 2 skip numeric, 3 skip any
  while (*p) {
    while (*p && isspace((unsigned char)*p)) p++;
    if (!*p) break;
    const char *t = p;
    while (*p && !isspace((unsigned char)*p)) p++;
    std::string tok(t, p);
    if (tok.size() < 2 || tok[0] != '%')
      return ing_err(g, PSH_ERR_ARG, "unsupported formatter token `%s'", tok.c_str());
    const bool skip = tok[1] == '*';
    std::string conv = tok.substr(skip ? 2 : 1);
    while (!conv.empty() && isdigit((unsigned char)conv[0])) conv.erase(0, 1);   // width
    int k;
    if (conv == "lf" || conv == "le" || conv == "lg" || conv == "lE" || conv == "lG" ||
        conv == "lF" || (skip && (conv == "f" || conv == "e" || conv == "g")))
      k = skip ? 2 : 0;
    else if (conv == "d" || conv == "i" || conv == "ld" || conv == "li" || conv == "u" ||
             conv == "lu")
      k = skip ? 2 : 1;
    else if (skip && (conv == "s" || conv == "c" || conv[0] == '['))
      k = 3;
    else
      return ing_err(g, PSH_ERR_ARG, "unsupported formatter token `%s'", tok.c_str());
    kind.push_back(k);
  }
  // value columns ($n counts the conversions that are not suppressed)
  std::vector<int> vfield;
  for (size_t i = 0; i < kind.size(); i++)
    if (kind[i] < 2) vfield.push_back((int)i);
  a.comment = (unsigned char)f->comment;
  if (ex) return compile_exprs(g, kind, vfield, ex, a, *prog);
  const int want[6] = {f->pos[0], f->pos[1], f->pos[2], f->wcomp, f->wfkp, f->nz};
  int last = -1;
  a.has = 0;
  for (int c = 0; c < 6; c++) {
    if (c < 3 && want[c] < 1)
      return ing_err(g, PSH_ERR_ARG, "position column %d is not given", c + 1);
    if (want[c] < 0 || want[c] > (int)vfield.size())
      return ing_err(g, PSH_ERR_ARG, "column $%d is beyond the formatter's %zu columns",
                     want[c], vfield.size());
    if (want[c] > 0) {
      a.has |= 1 << c;
      if (vfield[want[c] - 1] > last) last = vfield[want[c] - 1];
    }
  }
  a.nf = last + 1;   // conversions after the last used column are not read
  if (a.nf > kMaxFields) return ing_err(g, PSH_ERR_ARG, "more than %d formatter fields", kMaxFields);
  for (int i = 0; i < a.nf; i++) {
    a.map[i] = kind[i] == 3 ? F_SKIP_ANY : (kind[i] == 2 ? F_SKIP_NUM : F_UNUSED);
    a.integer[i] = kind[i] == 1;
  }
  for (int c = 0; c < 6; c++)
    if (want[c] > 0) {
      const int fi = vfield[want[c] - 1];
      if (a.map[fi] >= 0)
        return ing_err(g, PSH_ERR_ARG, "column $%d is used for two quantities", want[c]);
      a.map[fi] = (signed char)c;
    }
  a.expr = 0;
  return PSH_OK;
}

static int ingest_ascii(psh_ingest *g, const char *path, const psh_ascii_fmt *fmt,
                        const psh_ascii_expr *ex, int sim, psh_data *d_out, size_t cap) {
  if (!g) return PSH_ERR_ARG;
  if (!path || !fmt || !fmt->fmtr || !d_out) return ing_err(g, PSH_ERR_ARG, "null argument");
  if (g->busy) return ing_err(g, PSH_ERR_ARG, "an ingest is already in flight on this object");
  AsciiJob job;
  if (parse_fmtr(g, fmt, job.args, ex, &job.prog)) return g->rc;
  if (fmt->skip < 0) return ing_err(g, PSH_ERR_ARG, "negative number of header lines");
  job.skip = fmt->skip;
  job.path = path;
  job.fd = open(path, O_RDONLY);
  if (job.fd < 0) return ing_err(g, PSH_ERR_FILE, "cannot open the catalogue `%s'", path);
  g->rc = PSH_OK;
  g->err.clear();
  g->res = psh_ingest_result{};
  g->busy = true;
  double *out = reinterpret_cast<double *>(d_out);
  g->worker = std::thread([g, job, sim, out, cap]() {
    const int rc = run_ascii(g, job, sim, out, cap);
    if (rc && g->rc == PSH_OK) g->rc = rc;
    close(job.fd);
  });
  return PSH_OK;
}

extern "C" int psh_ingest_ascii(psh_ingest *g, const char *path, const psh_ascii_fmt *fmt,
                                int sim, psh_data *d_out, size_t cap) {
  return ingest_ascii(g, path, fmt, nullptr, sim, d_out, cap);
}

extern "C" int psh_ingest_ascii_expr(psh_ingest *g, const char *path, const psh_ascii_fmt *fmt,
                                     const psh_ascii_expr *ex, int sim, psh_data *d_out,
                                     size_t cap) {
  if (!g) return PSH_ERR_ARG;
  if (!ex) return ing_err(g, PSH_ERR_ARG, "null argument");
  return ingest_ascii(g, path, fmt, ex, sim, d_out, cap);
}

extern "C" int psh_parse_double(const char *s, size_t len, double *out) {
  if (!s || !out) return PSH_ERR_ARG;
  return pshdec::parse_double(s, s + len, out) ? PSH_OK : PSH_ERR_ARG;
}
