// psh_api.hip -- host orchestration of the mesh -> FFT -> P_l(k) path and
// the C-ABI declared in include/powspec_hip.h.
//
// The host side restates the reference's scalar set-up code (box definition
// src/genr_mesh.c:509-578, k bins src/multipole.c:306-394, alias factors
// src/multipole.c:46-100, normalisation src/multipole.c:1044-1162) and drives
// the gfx950 kernels of k_assign.hip, k_fft.hip and k_bin.hip on one HIP
// stream per context.
#include "../../include/powspec_hip.h"
#include "psh_internal.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#define PSH_BOX_CEIL 10   // src/define.h:95

namespace {

std::string g_last_error;

// Fault injection for tests of the error paths (no counterpart in the
// reference, whose allocation failures exit(), src/multipole.c:121-126):
// g_fault_alloc > 0 counts down the device allocations of contexts; the one
// that brings it to 0 fails as if the device were out of memory.  A negative
// value makes every allocation fail; 0 disables.  g_fault_graph != 0 makes
// graph instantiation fail.  Set from PSH_FAULT_ALLOC / PSH_FAULT_GRAPH at
// load or with psh_fault_inject.
std::atomic<long> g_fault_alloc{0};
std::atomic<int> g_fault_graph{0};
std::once_flag g_fault_env_once;

void fault_env_init() {
  std::call_once(g_fault_env_once, [] {
    if (const char *e = getenv("PSH_FAULT_ALLOC")) g_fault_alloc = atol(e);
    if (const char *e = getenv("PSH_FAULT_GRAPH")) g_fault_graph = atoi(e) != 0;
  });
}

bool fault_alloc_now() {
  fault_env_init();
  long v = g_fault_alloc.load();
  while (v > 0) {
    if (g_fault_alloc.compare_exchange_weak(v, v - 1)) return v == 1;
  }
  return v < 0;
}

bool fault_graph_now() {
  fault_env_init();
  return g_fault_graph.load() != 0;
}

// Stage events of one run (recorded only when profiling is on).
enum {
  E_START = 0,    // run entry
  E_ASSIGN,       // bucket kernels start
  E_DEPOSIT,      // deposit kernel start
  E_ASSIGNED,     // meshes complete
  E_ZDONE,        // z passes done
  E_YDONE,        // y passes done
  E_XDONE,        // fused x pass + binning done
  E_END,          // results reduced
  E_COUNT
};

struct Timer {
  hipEvent_t ev[E_COUNT] = {};
  // survey fields l > 0, three per (l, m): z pass start, y pass end, x pass end
  std::vector<hipEvent_t> lev;
};

}  // namespace

struct psh_ctx {
  psh_conf conf{};
  int dev = 0;
  // Device memory of an allocation-free run: every buffer of the context is
  // carved from this region (a bump allocator, 256-byte aligned; freeing is
  // a no-op) instead of hipMalloc'd.  nullptr: ordinary allocations.
  char *arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  int Ng = 0, M = 0, H = 0;
  size_t mstride = 0;            // floats per mesh
  int nmock_max = 1;
  hipStream_t stream = nullptr;
  bool survey = false;
  int ellmax = 0;
  int nfs = 1;   // survey l > 0 fields the S buffer holds (a pole's group)

  float *meshA = nullptr, *meshB = nullptr, *meshS = nullptr;
  int tile_req[3] = {0, 0, 0};   // deposit tiles set by psh_ctx_set_tiles / tuning
  int ysched = 0;                // schedule of the forward y passes (tuning)

  // assignment scratch
  PshAssignGeom ag{};
  size_t ntiles_mock = 0;
  uint32_t *counts = nullptr, *starts = nullptr, *cursors = nullptr;
  uint32_t *whist = nullptr;   // per-workgroup tile histograms of the count pass
  size_t whist_cap = 0, whist_words = 0;
  uint32_t *scan_tmp = nullptr;
  size_t tiles_cap = 0;
  PshDevPart *d_parts = nullptr;
  int parts_cap = 0;
  void *records = nullptr;
  size_t rec_cap = 0;
  int *d_flags = nullptr;

  // FFT tables
  float2 *tw = nullptr, *phase = nullptr;

  // geometry
  bool box_set = false;
  double L[3] = {0, 0, 0}, bmin[3] = {0, 0, 0}, cmax[3] = {0, 0, 0};
  int nbin = 0;
  std::vector<double> kedge, kcen;
  PshBins bins{};
  double *d_tab = nullptr;
  int *d_ylive = nullptr;   // per x-pass k_z tile, the largest live |k_y| (nullptr: all)
  int kzlim = 1 << 30;      // last k_z column of a live x-pass tile (z pass stores up to it)
  PshAxisTables ax{};
  bool geom_valid = false;
  std::vector<unsigned long long> cnt;
  std::vector<double> km, lcnt;
  unsigned long long *d_gcnt = nullptr;
  double *d_gkm = nullptr, *d_glcnt = nullptr;
  // bins per launch of the fused x pass (its LDS holds a window of the bins)
  int xwin_sim = 0, xwin_survey = 0;
  // CUs of the device and workgroups per CU of the x pass kernels (boxes;
  // survey l = 0 and l > 0 fields) at their first, largest window
  int ncu = 256;
  int xres_sim = 1, xres_s0 = 1, xres_sl = 1;

  // |mu| cells of box runs and the learned alias correction of plain boxes
  int nmu = 1;
  std::vector<double> cells;          // raw sums of the last box run [nl][nbin][nmu]
  double last_norm = 0, last_shot = 0;
  float *d_mlp = nullptr;             // [mlp_nnet][psh_mlp_param_count()]
  int mlp_nnet = 0;
  float *d_ax = nullptr, *d_ay = nullptr;   // feature rows, network outputs
  double *d_acorr = nullptr;          // corrected raw sums [nl][nbin]
  double *d_afeat = nullptr;          // feature scratch (per-bin levels, band summary)
  int afeat_nbin = 0;                 // bins d_afeat holds
  size_t arows_cap = 0;

  // binning
  double *partials = nullptr;
  size_t partials_cap = 0;
  double *d_bins = nullptr;
  size_t bins_cap = 0;
  double *d_bounds = nullptr;

  // device particle staging (host-catalogue runs)
  double *d_cat = nullptr;
  size_t d_cat_cap = 0;

  // timing
  bool profiling = false;
  Timer tm;
  double ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};

  // graph of a batched run
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_nmock = 0;
  size_t graph_ntotal = 0;
  const void *graph_data = nullptr;

  std::string err;
};

struct psh_mesh {
  psh_ctx *ctx = nullptr;
  int Ng = 0;
  double bsize[3] = {0, 0, 0};
  double bmin[3] = {0, 0, 0};
  int assign = 0, intlace = 0;
  int num = 0;
  bool assigned[2] = {false, false};
  // per-catalogue device particles kept between genr_mesh and powspec
  double *d_data[2] = {nullptr, nullptr};
  size_t ndata[2] = {0, 0};
  double *d_rand[2] = {nullptr, nullptr};
  size_t nrand[2] = {0, 0};
  double sw_d[2] = {0, 0}, sw_r[2] = {0, 0};   // weight sums of the catalogues
};

namespace {

int set_err(psh_ctx *ctx, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  g_last_error = buf;
  return code;
}

#define HIPCHK(ctx, call)                                                        \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      return set_err(ctx, PSH_ERR_DEVICE, "%s: %s (%s:%d)", #call,               \
                     hipGetErrorString(e_), __FILE__, __LINE__);                 \
  } while (0)

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// Deposit tile shapes the kernels accept: powers of two, at least the
// window width W = scheme + 1 along every axis (a window plus the shifted
// grid's extra cell then meets at most two tiles per axis, which the
// bucketing assumes), rows of at least 4 cells along z, LDS of both grids
// (z rows padded for the window) within 160 KB at `cell_bytes` per cell
// (4: 32-bit fixed point of boxes, 8: 64-bit of surveys).
bool tiles_ok(const int t[3], int scheme, int intlace, int cell_bytes) {
  for (int a = 0; a < 3; a++)
    if (!is_pow2(t[a]) || t[a] > 2048 || t[a] < scheme + 1) return false;
  return t[2] >= 4 &&
         psh_tile_cells(t, scheme) * cell_bytes * (intlace ? 2 : 1) <= 160 * 1024;
}

// src/multipole.c:46-100: per-axis alias correction, series branch for
// pik <= 0.01 (which includes every negative frequency).
double alias_corr(int intlace, int assign, double pik) {
  double fac;
  if (intlace) {
    if (pik > 0.01) {
      fac = pik / sin(pik);
    } else {
      double t = pik * pik;
      fac = 1 + (1.0 / 6) * t + (7.0 / 360) * t * t + (31.0 / 15120) * t * t * t;
    }
    fac *= fac;
    switch (assign) {
      case PSH_ASSIGN_NGP: return fac;
      case PSH_ASSIGN_CIC: return fac * fac;
      case PSH_ASSIGN_TSC: return fac * fac * fac;
      default: fac *= fac; return fac * fac;
    }
  }
  if (pik > 0.01) {
    fac = sin(pik);
  } else {
    double t = pik * pik;
    fac = (1 - (1.0 / 6) * t + (1.0 / 120) * t * t - (1.0 / 5040) * t * t * t) * pik;
  }
  fac *= fac;
  switch (assign) {
    case PSH_ASSIGN_NGP: return 1;
    case PSH_ASSIGN_CIC: return 1 / (1 - (2.0 / 3) * fac);
    case PSH_ASSIGN_TSC: return 1 / (1 - fac + (2.0 / 15) * fac * fac);
    default:
      return 1 / (1 - (4.0 / 3) * fac + 0.4 * fac * fac - (4.0 / 315) * fac * fac * fac);
  }
}

// Normalisation of the real-form Y_lm with the associated Legendre function
// without Condon-Shortley phase.
float ylm_norm(int ell, int m) {
  int am = m < 0 ? -m : m;
  double ratio = 1;
  for (int i = ell - am + 1; i <= ell + am; i++) ratio /= i;
  double K = sqrt((2 * ell + 1) / (4 * M_PI) * ratio);
  if (m != 0) K *= sqrt(2.0);
  return (float)K;
}

int validate_conf(const psh_conf *c, psh_ctx *ctx) {
  if (!c) return set_err(ctx, PSH_ERR_ARG, "configuration not given");
  if (c->ndata < 1 || c->ndata > 2)
    return set_err(ctx, PSH_ERR_ARG, "invalid number of catalogues: %d", c->ndata);
  if (c->gsize <= 1 || c->gsize > 65536)
    return set_err(ctx, PSH_ERR_ARG, "invalid GRID_SIZE: %d", c->gsize);
  if (!psh_grid_supported(c->gsize))
    return set_err(ctx, PSH_ERR_UNSUPPORTED,
                   "GRID_SIZE %d: this build supports powers of two from 4 to 2048 "
                   "and 3 * 2^k (48 to 1536) or 5 * 2^k (80 to 1280)",
                   c->gsize);
  if (c->assign < 0 || c->assign > 3)
    return set_err(ctx, PSH_ERR_ARG, "invalid PARTICLE_ASSIGN: %d", c->assign);
  if (c->npole < 1 || c->npole > PSH_MAX_ELL + 1)
    return set_err(ctx, PSH_ERR_ARG, "invalid number of multipoles: %d", c->npole);
  for (int i = 0; i < c->npole; i++) {
    if (c->poles[i] < 0 || c->poles[i] > PSH_MAX_ELL)
      return set_err(ctx, PSH_ERR_ARG, "MULTIPOLE must be between 0 and %d", PSH_MAX_ELL);
    if (i && c->poles[i] <= c->poles[i - 1])
      return set_err(ctx, PSH_ERR_ARG, "MULTIPOLE must be sorted and unique");
  }
  if (!(c->kbin > 0)) return set_err(ctx, PSH_ERR_ARG, "BIN_SIZE must be > 0");
  if (c->issim) {
    for (int a = 0; a < 3; a++)
      if (!(c->bsize[a] > 0)) return set_err(ctx, PSH_ERR_ARG, "BOX_SIZE must be > 0");
  }
  if (c->iscross && c->ndata != 2)
    return set_err(ctx, PSH_ERR_ARG, "cross power spectra need two catalogues");
  return PSH_OK;
}

// Every device buffer of a context comes from ctx_malloc and goes back with
// ctx_free: hipMalloc/hipFree, or the context's arena.
int ctx_malloc(psh_ctx *ctx, void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 1;
  if (ctx && ctx->arena) {
    const size_t off = (ctx->arena_used + 255) & ~(size_t)255;
    if (off > ctx->arena_cap || bytes > ctx->arena_cap - off)
      return set_err(ctx, PSH_ERR_MEMORY,
                     "the run needs more than the %zu bytes of the static device scratch "
                     "(%zu in use, %zu more requested)",
                     ctx->arena_cap, off, bytes);
    *p = ctx->arena + off;
    ctx->arena_used = off + bytes;
    return PSH_OK;
  }
  if (fault_alloc_now())
    return set_err(ctx, PSH_ERR_MEMORY,
                   "device allocation of %zu bytes failed: injected fault", bytes);
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_err(ctx, PSH_ERR_MEMORY, "device allocation of %zu bytes failed: %s", bytes,
                   hipGetErrorString(e));
  }
  return PSH_OK;
}

void ctx_free(psh_ctx *ctx, void *p) {
  if (!p || (ctx && ctx->arena)) return;
  (void)hipFree(p);
}

template <class T>
int dev_alloc(psh_ctx *ctx, T **p, size_t count) {
  if (*p) {
    ctx_free(ctx, *p);
    *p = nullptr;
  }
  return ctx_malloc(ctx, (void **)p, sizeof(T) * count);
}

// The S buffers of the survey l > 0 fields: one group of up to 2 l_max + 1
// fields, as many as half of the free device memory holds (at least one).
// PSH_SURVEY_FIELDS caps the group (tests of poles split over groups).
int alloc_survey_fields(psh_ctx *ctx) {
  const size_t field = sizeof(float) * ctx->mstride;
  int nfs = std::min(2 * ctx->ellmax + 1, PSH_MAX_YLM_FIELDS);
  if (const char *e = getenv("PSH_SURVEY_FIELDS")) {
    const int cap = atoi(e);
    if (cap >= 1 && cap < nfs) nfs = cap;
  }
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
    const size_t fit = free_b / 2 / field;
    if ((size_t)nfs > fit) nfs = (int)std::max<size_t>(1, fit);
  } else {
    (void)hipGetLastError();
    nfs = 1;
  }
  ctx->nfs = nfs;
  return dev_alloc(ctx, &ctx->meshS, ctx->mstride * (size_t)nfs);
}

// k bins: src/multipole.c:334-394.
int make_bins(psh_ctx *ctx) {
  const psh_conf &c = ctx->conf;
  double bmax = ctx->L[0] > ctx->L[1] ? ctx->L[0] : ctx->L[1];
  if (bmax < ctx->L[2]) bmax = ctx->L[2];
  double kny = M_PI * ctx->Ng / bmax;
  if (c.logscale) kny = log10(kny);
  double kmax = kny;
  if (c.kmax > 0 && kmax > c.kmax) kmax = c.kmax;
  double nb = round((kmax - c.kmin) / c.kbin);
  if (nb >= 2147483647.0)
    return set_err(ctx, PSH_ERR_PK, "too many wave number bins due to the small bin size");
  int nbin = (int)nb;
  if (c.kmin + c.kbin * nbin > kny) nbin -= 1;
  if (nbin < 1)
    return set_err(ctx, PSH_ERR_PK,
                   "not enough k bins given the Nyquist frequency and the bin size");
  ctx->nbin = nbin;
  ctx->kedge.assign(nbin + 1, 0.0);
  ctx->kcen.assign(nbin, 0.0);
  for (int i = 0; i <= nbin; i++) ctx->kedge[i] = c.kmin + c.kbin * i;
  for (int i = 0; i < nbin; i++) ctx->kcen[i] = (ctx->kedge[i] + ctx->kedge[i + 1]) * 0.5;
  PshBins &b = ctx->bins;
  b.nbin = nbin;
  b.logscale = c.logscale;
  b.kedge0 = ctx->kedge[0];
  b.kedgeN = ctx->kedge[nbin];
  b.dk = c.kbin;
  b.nl = c.npole;
  for (int l = 0; l < PSH_MAX_NL; l++) b.poles[l] = l < c.npole ? c.poles[l] : 0;
  for (int a = 0; a < 3; a++) {
    b.los[a] = c.issim ? c.los[a] : 0.0;
    b.losf[a] = (float)b.los[a];
  }
  b.kedge0f = (float)b.kedge0;
  b.invdkf = (float)(1.0 / c.kbin);
  b.k0dk = (float)fabs(b.kedge0 / c.kbin);
  b.nmu = c.issim ? ctx->nmu : 1;
  // the fused x pass keeps each workgroup's bin sums in LDS next to its
  // tables and exchange tile (simulation boxes: npole sums per bin; survey
  // fields: one); more bins than fit are binned in windows of bins, one
  // launch each (the reference has no bin limit, src/multipole.c:340-356)
  ctx->xwin_sim = psh_xbin_window_bins(ctx->Ng, c.intlace != 0, c.npole * b.nmu);
  ctx->xwin_survey = std::min(psh_xbin_window_bins(ctx->Ng, c.intlace != 0, 1),
                              psh_xbin_window_bins(ctx->Ng, false, 1));
  if ((c.issim ? ctx->xwin_sim : ctx->xwin_survey) < 1)
    return set_err(ctx, PSH_ERR_UNSUPPORTED,
                   "the fused x pass has no LDS left for bin sums at GRID_SIZE %d", ctx->Ng);
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      ncu > 0)
    ctx->ncu = ncu;
  auto res = [&](int epi, bool il, int nacc) {
    const int r = psh_xbin_per_cu(ctx->Ng, epi, il, nacc);
    return r > 0 ? r : 1;
  };
  if (c.issim) {
    ctx->xres_sim = res(PSH_EPI_SIM, c.intlace != 0,
                        std::min(ctx->xwin_sim, b.nbin) * c.npole * b.nmu);
  } else {
    ctx->xres_s0 = res(PSH_EPI_SURVEY0, c.intlace != 0, std::min(ctx->xwin_survey, b.nbin));
    ctx->xres_sl = res(PSH_EPI_SURVEYL, false, std::min(ctx->xwin_survey, b.nbin));
  }
  return PSH_OK;
}

// Per-axis FP64 tables (signed frequency index for all three axes).
int make_tables(psh_ctx *ctx) {
  const int N = ctx->Ng;
  std::vector<double> h((size_t)12 * N);
  const double fac = M_PI / N;
  for (int a = 0; a < 3; a++) {
    const double vec = 2 * M_PI / ctx->L[a];
    const double dfac = 1 / ctx->L[a];
    double *kv = h.data() + (size_t)(4 * a + 0) * N;
    double *ksq = h.data() + (size_t)(4 * a + 1) * N;
    double *al = h.data() + (size_t)(4 * a + 2) * N;
    double *kd = h.data() + (size_t)(4 * a + 3) * N;
    for (int i = 0; i < N; i++) {
      const int s = (i <= (N >> 1)) ? i : i - N;
      double ki = s;
      al[i] = alias_corr(ctx->conf.intlace, ctx->conf.assign, ki * fac);
      ki *= vec;
      kv[i] = ki;
      ksq[i] = ki * ki;
      kd[i] = dfac * s;
    }
  }
  // FP32 roundings of kv, ksq, al appended after the FP64 tables
  std::vector<float> hf((size_t)9 * N);
  for (int a = 0; a < 3; a++)
    for (int t = 0; t < 3; t++)
      for (int i = 0; i < N; i++)
        hf[(size_t)(3 * a + t) * N + i] = (float)h[(size_t)(4 * a + t) * N + i];
  const size_t nd = h.size() + (hf.size() + 1) / 2;
  int rc = dev_alloc(ctx, &ctx->d_tab, nd);
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpy(ctx->d_tab, h.data(), sizeof(double) * h.size(),
                        hipMemcpyHostToDevice));
  float *dtf = reinterpret_cast<float *>(ctx->d_tab + h.size());
  HIPCHK(ctx, hipMemcpy(dtf, hf.data(), sizeof(float) * hf.size(), hipMemcpyHostToDevice));
  for (int a = 0; a < 3; a++) {
    float amax = 0.f;
    for (int i = 0; i < N; i++) amax = std::max(amax, hf[(size_t)(3 * a + 2) * N + i]);
    ctx->ax.almax[a] = amax;
    ctx->ax.kv[a] = ctx->d_tab + (size_t)(4 * a + 0) * N;
    ctx->ax.ksq[a] = ctx->d_tab + (size_t)(4 * a + 1) * N;
    ctx->ax.al[a] = ctx->d_tab + (size_t)(4 * a + 2) * N;
    ctx->ax.kd[a] = ctx->d_tab + (size_t)(4 * a + 3) * N;
    ctx->ax.kvf[a] = dtf + (size_t)(3 * a + 0) * N;
    ctx->ax.ksqf[a] = dtf + (size_t)(3 * a + 1) * N;
    ctx->ax.alf[a] = dtf + (size_t)(3 * a + 2) * N;
  }
  // Live k_y bound per tile of the fused x pass (k_z columns [q CX, q CX +
  // CX)): the largest |k_y| index for which the tile's nearest mode (k_x = 0,
  // k_z = q CX) lies before the last bin edge, with the binning's own FP64
  // |k| = sqrt((k_x^2 + k_y^2) + k_z^2) (src/multipole.c:147-158), which
  // only grows with |k_x|, |k_y| and k_z.  The y pass stores no output past
  // it and the x pass skips the tiles; -1 when no k_y is live.  Log bins
  // keep every tile.
  ctx->kzlim = 1 << 30;
  if (ctx->conf.logscale) {
    ctx_free(ctx, ctx->d_ylive);
    ctx->d_ylive = nullptr;
  } else {
    const int CX = psh_xbin_tile_cols(N);
    const int ntzx = (N / 2 + 1 + CX - 1) / CX;
    const double *ksqy = h.data() + (size_t)(4 * 1 + 1) * N;
    const double *ksqz = h.data() + (size_t)(4 * 2 + 1) * N;
    std::vector<int> yl(ntzx, -1);
    for (int q = 0; q < ntzx; q++) {
      const int kz0 = q * CX;
      for (int ky = 0; ky <= N / 2; ky++) {
        double m2 = ksqy[ky];
        if (kz0 > 0) m2 = m2 + ksqz[kz0];
        if (!(sqrt(m2) < ctx->bins.kedgeN)) break;
        yl[q] = ky;
      }
    }
    // the z pass stores k_z up to the last column of the last live tile
    ctx->kzlim = -1;
    for (int q = 0; q < ntzx; q++)
      if (yl[q] >= 0) ctx->kzlim = std::min(N / 2, (q + 1) * CX - 1);
    if ((rc = dev_alloc(ctx, &ctx->d_ylive, yl.size()))) return rc;
    HIPCHK(ctx, hipMemcpy(ctx->d_ylive, yl.data(), sizeof(int) * yl.size(),
                          hipMemcpyHostToDevice));
  }
  return PSH_OK;
}

int ensure_geometry(psh_ctx *ctx) {
  if (ctx->geom_valid) return PSH_OK;
  const int nb = ctx->nbin, nl = ctx->conf.npole;
  int rc;
  if ((rc = dev_alloc(ctx, &ctx->d_gcnt, (size_t)nb))) return rc;
  if ((rc = dev_alloc(ctx, &ctx->d_gkm, (size_t)nb))) return rc;
  if ((rc = dev_alloc(ctx, &ctx->d_glcnt, (size_t)nb * nl))) return rc;
  // per-workgroup partials of the geometry pass, freed when it is done
  char *gscratch = nullptr;
  if ((rc = dev_alloc(ctx, &gscratch, psh_geometry_scratch_bytes(ctx->bins)))) return rc;
  hipError_t gerr = psh_geometry(ctx->Ng, ctx->ax, ctx->bins, ctx->d_gcnt, ctx->d_gkm,
                                 ctx->d_glcnt, gscratch, ctx->stream);
  if (gerr == hipSuccess) gerr = hipStreamSynchronize(ctx->stream);
  ctx_free(ctx, gscratch);
  HIPCHK(ctx, gerr);
  ctx->cnt.assign(nb, 0);
  ctx->km.assign(nb, 0.0);
  ctx->lcnt.assign((size_t)nb * nl, 0.0);
  HIPCHK(ctx, hipMemcpyAsync(ctx->cnt.data(), ctx->d_gcnt, sizeof(unsigned long long) * nb,
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->km.data(), ctx->d_gkm, sizeof(double) * nb,
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->lcnt.data(), ctx->d_glcnt, sizeof(double) * nb * nl,
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < nb; i++)
    if (ctx->cnt[i]) ctx->km[i] /= (double)ctx->cnt[i];
  ctx->geom_valid = true;
  return PSH_OK;
}

// Assignment geometry of the context's box and deposit tiles.
void set_assign_geom(psh_ctx *ctx) {
  PshAssignGeom &g = ctx->ag;
  g.Ng = ctx->Ng;
  g.scheme = ctx->conf.assign;
  g.intlace = ctx->conf.intlace;
  g.check_box = ctx->conf.issim ? 1 : 0;   // simulation boxes: [0, L) per axis
  // 8x16x64 cells per grid by default (measured on MI355X against 8x16x32:
  // 2-10 % faster for every scheme, grid size and interlacing timed; it fits
  // the LDS for every scheme with 64-bit survey cells too);
  // PSH_TILE="tx,ty,tz" overrides (powers of two, tz >= 4), and so do
  // psh_ctx_set_tiles and a tuning cache
  int tl[3] = {8, 16, 64};
  if (const char *env = getenv("PSH_TILE")) {
    int t[3] = {0, 0, 0};
    if (sscanf(env, "%d,%d,%d", &t[0], &t[1], &t[2]) == 3 &&
        tiles_ok(t, ctx->conf.assign, ctx->conf.intlace, ctx->survey ? 8 : 4))
      for (int a = 0; a < 3; a++) tl[a] = t[a];
  }
  if (ctx->tile_req[0])
    for (int a = 0; a < 3; a++) tl[a] = ctx->tile_req[a];
  // tiles are powers of two that divide the grid: at most its largest
  // power-of-two factor (the grid itself when it is a power of two)
  const int p2 = ctx->Ng & -ctx->Ng;
  for (int a = 0; a < 3; a++) g.tile[a] = p2 < tl[a] ? p2 : tl[a];
  for (int a = 0; a < 3; a++) g.ntile[a] = ctx->Ng / g.tile[a];
  for (int a = 0; a < 3; a++) {
    g.tile_shift[a] = 0;
    while ((1 << g.tile_shift[a]) < g.tile[a]) g.tile_shift[a]++;
  }
  g.row_pitch = 2 * ctx->H;
  for (int a = 0; a < 3; a++) {
    g.bmin[a] = ctx->bmin[a];
    g.smin[a] = ctx->bmin[a] - 0.5 * ctx->L[a] / ctx->Ng;   // src/genr_mesh.c:817-818
    g.L[a] = ctx->L[a];
  }
  // tiles of whole z rows in the fused shape: the deposit transforms its rows
  // (boxes; surveys keep the real mesh for their Y_lm-weighted passes)
  int zt[3];
  g.zfuse = !ctx->survey && psh_zfuse_tile(ctx->Ng, ctx->conf.intlace, zt) &&
            g.tile[0] == zt[0] && g.tile[1] == zt[1] && g.tile[2] == zt[2];
  g.kzlim = ctx->kzlim;
  g.tw = ctx->tw;
  ctx->ntiles_mock = (size_t)g.ntile[0] * g.ntile[1] * g.ntile[2];
}

void drop_graph(psh_ctx *ctx);

int set_box(psh_ctx *ctx, const double L[3], const double bmin[3]) {
  bool same = ctx->box_set;
  for (int a = 0; a < 3; a++)
    if (ctx->L[a] != L[a] || ctx->bmin[a] != bmin[a]) same = false;
  if (same) return PSH_OK;
  // a captured batch graph holds the tables and geometry of the old box,
  // which are about to be reallocated; until the new box is complete the
  // context has none (a failed set-up is redone in full by the next call)
  drop_graph(ctx);
  ctx->box_set = false;
  ctx->geom_valid = false;
  for (int a = 0; a < 3; a++) {
    ctx->L[a] = L[a];
    ctx->bmin[a] = bmin[a];
  }
  int rc;
  if ((rc = make_bins(ctx))) return rc;
  if ((rc = make_tables(ctx))) return rc;
  set_assign_geom(ctx);
  ctx->box_set = true;
  return PSH_OK;
}

// def_box from coordinate bounds (src/genr_mesh.c:509-578).
int def_box(psh_ctx *ctx, const double mn[3], const double mx[3]) {
  const psh_conf &c = ctx->conf;
  const char ax[3] = {'x', 'y', 'z'};
  double L[3], bmin[3];
  for (int i = 0; i < 3; i++) {
    if (mn[i] > mx[i])
      return set_err(ctx, PSH_ERR_CATA, "invalid %c coordinate value in the catalogs", ax[i]);
    if (c.issim) {
      if (mn[i] < 0) return set_err(ctx, PSH_ERR_CATA, "%c coordinate below 0: %lf", ax[i], mn[i]);
      if (mx[i] >= c.bsize[i])
        return set_err(ctx, PSH_ERR_CATA, "%c coordinate not smaller than BOX_SIZE: %lf",
                       ax[i], mx[i]);
    }
  }
  if (c.issim) {
    for (int i = 0; i < 3; i++) {
      L[i] = c.bsize[i];
      bmin[i] = 0;
    }
  } else if (c.has_bsize) {
    for (int i = 0; i < 3; i++) {
      if (mx[i] - mn[i] > c.bsize[i])
        return set_err(ctx, PSH_ERR_CATA,
                       "BOX_SIZE is too small for the %c coordinates, should be at least %.10lg",
                       ax[i], mx[i] - mn[i]);
      bmin[i] = (mx[i] + mn[i] - c.bsize[i]) * 0.5;
      L[i] = c.bsize[i];
    }
  } else {
    for (int i = 0; i < 3; i++) {
      L[i] = ceil((mx[i] - mn[i]) * (1 + c.bpad[i]) / PSH_BOX_CEIL) * PSH_BOX_CEIL;
      bmin[i] = (mx[i] + mn[i] - L[i]) * 0.5;
    }
  }
  for (int i = 0; i < 3; i++) ctx->cmax[i] = mx[i];
  return set_box(ctx, L, bmin);
}

// A captured graph holds the scratch pointers of its capture: any
// reallocation drops it (it is captured again on the next batched call).
void drop_graph(psh_ctx *ctx) {
  if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->gexec = nullptr;
  ctx->graph = nullptr;
}

int ensure_assign_scratch(psh_ctx *ctx, int nmock, int nparts, size_t ntotal) {
  int rc;
  const size_t ntiles = ctx->ntiles_mock * nmock;
  // each capacity is cleared before its buffers are replaced and set only
  // once they all exist, so a failed allocation leaves no stale capacity
  if (ntiles + 1 > ctx->tiles_cap) {
    drop_graph(ctx);
    ctx->tiles_cap = 0;
    if ((rc = dev_alloc(ctx, &ctx->counts, ntiles + 1))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->starts, ntiles + 1))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->cursors, ntiles + 1))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->scan_tmp, psh_assign_scan_bytes(ntiles) / 4 + 64)))
      return rc;
    ctx->tiles_cap = ntiles + 1;
  }
  if (nparts > ctx->parts_cap) {
    drop_graph(ctx);
    ctx->parts_cap = 0;
    if ((rc = dev_alloc(ctx, &ctx->d_parts, (size_t)nparts))) return rc;
    ctx->parts_cap = nparts;
  }
  if (!ctx->d_flags) {
    if ((rc = dev_alloc(ctx, &ctx->d_flags, 4))) return rc;
  }
  ctx->whist_words = psh_assign_whist_words(ctx->ag, ntotal);
  if (ctx->whist_words > ctx->whist_cap) {
    drop_graph(ctx);
    ctx_free(ctx, ctx->whist);
    ctx->whist = nullptr;
    ctx->whist_cap = 0;
    // optional: without the histograms the fill pass recounts
    const std::string keep = ctx->err;
    if (ctx_malloc(ctx, (void **)&ctx->whist, sizeof(uint32_t) * ctx->whist_words) == PSH_OK) {
      ctx->whist_cap = ctx->whist_words;
    } else {
      ctx->whist_words = 0;
      ctx->err = keep;
    }
  }
  // expected (particle, tile) pairs: prod_a (1 + (W' - 1) / T_a), W' the
  // union window; 1.5x margin, grown on overflow
  const int W = ctx->conf.assign + 1 + (ctx->conf.intlace ? 1 : 0);
  double dup = 1.0;
  for (int a = 0; a < 3; a++) dup *= 1.0 + (double)(W - 1) / ctx->ag.tile[a];
  size_t need = (size_t)(ntotal * dup * 1.5) + 4096;
  if (need > ctx->rec_cap) {
    drop_graph(ctx);
    ctx_free(ctx, ctx->records);
    ctx->records = nullptr;
    ctx->rec_cap = 0;
    if ((rc = ctx_malloc(ctx, &ctx->records, need * psh_assign_record_bytes()))) return rc;
    ctx->rec_cap = need;
  }
  return PSH_OK;
}

int grow_records(psh_ctx *ctx) {
  drop_graph(ctx);
  const size_t need = ctx->rec_cap * 2;
  ctx_free(ctx, ctx->records);
  ctx->records = nullptr;
  ctx->rec_cap = 0;
  int rc;
  if ((rc = ctx_malloc(ctx, &ctx->records, need * psh_assign_record_bytes()))) return rc;
  ctx->rec_cap = need;
  return PSH_OK;
}

int ensure_partials(psh_ctx *ctx, int nmock, int grid, int nacc, int slots) {
  int rc;
  size_t need = (size_t)nmock * grid * nacc;
  // a captured batch graph holds these buffers: drop it before reallocating
  if (need > ctx->partials_cap) {
    drop_graph(ctx);
    ctx->partials_cap = 0;
    if ((rc = dev_alloc(ctx, &ctx->partials, need))) return rc;
    ctx->partials_cap = need;
  }
  size_t needb = (size_t)nmock * nacc * slots;
  if (needb > ctx->bins_cap) {
    drop_graph(ctx);
    ctx->bins_cap = 0;
    if ((rc = dev_alloc(ctx, &ctx->d_bins, needb))) return rc;
    ctx->bins_cap = needb;
  }
  return PSH_OK;
}

// Persistent workgroups per mock of the fused x pass: as many as the device
// holds at once (CUs x per_cu, the kernel's workgroups per CU), shared among
// the mocks, a multiple of 8 (one share per XCD), at most one per tile.
// More workgroups than that only add rounds of set-up, partial-sum rows and
// a ragged last round: measured on MI355X, 256 workgroups instead of 2048
// take the interlaced N_g = 1024 PCS box pass from 3.70 to 3.48 ms and 512
// the plain N_g = 512 pass with 8 |mu| cells from 0.43 to 0.36 ms; 1.5
// rounds (384 at one per CU) cost 20 %.
int xgrid_for(psh_ctx *ctx, int nmock, int per_cu) {
  const int C = 16;
  const int ntz = (ctx->M + 1 + C - 1) / C;
  const long ntiles = (long)ctx->Ng * ntz;
  long g = ((long)ctx->ncu * per_cu / nmock) & ~7L;
  if (g < 8) g = 8;
  if (g > ntiles) g = ntiles;
  return (int)g;
}

// A named host range around a stage's launches (roctx; shown by
// rocprofv3 --marker-trace next to the kernels it enqueued).  The calls cost
// a few nanoseconds when no profiler is attached.
struct Mark {
  explicit Mark(const char *name) { roctxRangePushA(name); }
  ~Mark() { roctxRangePop(); }
  Mark(const Mark &) = delete;
  Mark &operator=(const Mark &) = delete;
};

void ev_record(psh_ctx *ctx, int i) {
  if (ctx->profiling && ctx->tm.ev[i]) (void)hipEventRecord(ctx->tm.ev[i], ctx->stream);
}

// -1 for a pair of events the run did not record; the failed query's error
// is cleared, or the next launch check would report it
double ev_ms(psh_ctx *ctx, int a, int b) {
  float t;
  if (hipEventElapsedTime(&t, ctx->tm.ev[a], ctx->tm.ev[b]) == hipSuccess) return t;
  (void)hipGetLastError();
  return -1.0;
}

// ms: [0] assignment (incl. bounds), [1] z+y passes, [2] fused x pass,
// [3] total, [4] deposit kernel, [5] z passes, [6] y passes, [7] bucketing
void ev_collect(psh_ctx *ctx) {
  if (!ctx->profiling) return;
  ctx->ms[0] = ev_ms(ctx, E_START, E_ASSIGNED);
  ctx->ms[1] = ev_ms(ctx, E_ASSIGNED, E_YDONE);
  ctx->ms[2] = ev_ms(ctx, E_YDONE, E_XDONE);
  ctx->ms[3] = ev_ms(ctx, E_START, E_END);
  ctx->ms[4] = ev_ms(ctx, E_DEPOSIT, E_ASSIGNED);
  ctx->ms[5] = ev_ms(ctx, E_ASSIGNED, E_ZDONE);
  ctx->ms[6] = ev_ms(ctx, E_ZDONE, E_YDONE);
  ctx->ms[7] = ev_ms(ctx, E_ASSIGN, E_DEPOSIT);
}

// Assign parts (device table already built on host) into meshes.
int run_assign(psh_ctx *ctx, const std::vector<PshDevPart> &hp, int nmock,
               float *meshA, float *meshB) {
  const Mark mark("psh:assign");
  size_t ntotal = 0;
  ctx->ag.wide = ctx->ag.zfuse && psh_zfuse_wide(ctx->Ng);   // the fused shape's cells
  for (auto &p : hp) {
    ntotal += p.n;
    if (p.wscale < 0.f) ctx->ag.wide = 1;   // data - alpha * randoms
  }
  if (ctx->ag.wide && !tiles_ok(ctx->ag.tile, ctx->conf.assign, ctx->conf.intlace, 8))
    return set_err(ctx, PSH_ERR_UNSUPPORTED,
                   "tile %dx%dx%d too large for 64-bit accumulation of weights of both signs",
                   ctx->ag.tile[0], ctx->ag.tile[1], ctx->ag.tile[2]);
  int rc;
  if ((rc = ensure_assign_scratch(ctx, nmock, (int)hp.size(), ntotal))) return rc;
  for (int attempt = 0; attempt < 4; attempt++) {
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_parts, hp.data(), sizeof(PshDevPart) * hp.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, psh_fill_zero_u32((uint32_t *)ctx->d_flags, 4, ctx->stream));
    ev_record(ctx, E_ASSIGN);
    HIPCHK(ctx, psh_assign_run(ctx->ag, ctx->d_parts, (int)hp.size(), ntotal, nmock,
                               ctx->counts, ctx->starts, ctx->cursors, ctx->records,
                               ctx->rec_cap, meshA, meshB, ctx->mstride, ctx->scan_tmp,
                               ctx->d_flags, ctx->stream,
                               ctx->profiling ? ctx->tm.ev[E_DEPOSIT] : nullptr,
                               ctx->whist_words ? ctx->whist : nullptr));
    ev_record(ctx, E_ASSIGNED);
    int flags = 0;
    HIPCHK(ctx, hipMemcpyAsync(&flags, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (flags & 4)
      return set_err(ctx, PSH_ERR_CATA, "particle coordinate outside the box [0, BOX_SIZE)");
    if (flags & 2)
      return set_err(ctx, PSH_ERR_MESH, "inconsistent interlaced cell of a particle");
    if (!(flags & 1)) return PSH_OK;
    if ((rc = grow_records(ctx))) return rc;
  }
  return set_err(ctx, PSH_ERR_MESH, "assignment record buffer overflow");
}

// FFTs of the plain (and shifted) meshes of nmock mocks, fused binning.
int run_sim_fft_bin(psh_ctx *ctx, int nmock, double *out_bins /*[nmock][nacc]*/) {
  const Mark mark("psh:fft_bin");
  const int N = ctx->Ng;
  PshYlmR noy{};
  float2 *A = reinterpret_cast<float2 *>(ctx->meshA);
  float2 *B = ctx->conf.intlace ? reinterpret_cast<float2 *>(ctx->meshB) : nullptr;
  if (!ctx->ag.zfuse) {   // else the deposit wrote half-complex rows already
    HIPCHK(ctx, psh_fft_z_r2c(N, ctx->meshA, A, nmock, ctx->mstride, ctx->tw, noy,
                              ctx->stream, ctx->kzlim));
    if (B)
      HIPCHK(ctx, psh_fft_z_r2c(N, ctx->meshB, B, nmock, ctx->mstride, ctx->tw, noy,
                                ctx->stream, ctx->kzlim));
  }
  ev_record(ctx, E_ZDONE);
  HIPCHK(ctx, psh_fft_y(N, A, nmock, ctx->mstride, ctx->tw, 0, ctx->stream, ctx->d_ylive,
                        ctx->ysched));
  if (B)
    HIPCHK(ctx, psh_fft_y(N, B, nmock, ctx->mstride, ctx->tw, 0, ctx->stream, ctx->d_ylive,
                          ctx->ysched));
  ev_record(ctx, E_YDONE);
  const int grid = xgrid_for(ctx, nmock, ctx->xres_sim);
  const int nacc = psh_xbin_nacc(ctx->bins, PSH_EPI_SIM);
  int rc;
  if ((rc = ensure_partials(ctx, nmock, grid, nacc, 1))) return rc;
  PshXBinArgs a{};
  a.epi = PSH_EPI_SIM;
  a.A = A;
  a.B = B;
  a.nmock = nmock;
  a.mstride_floats = ctx->mstride;
  a.ax = ctx->ax;
  a.bins = ctx->bins;
  a.phase = ctx->phase;
  a.partials = ctx->partials;
  a.grid = grid;
  a.vec0f = (float)(2 * M_PI / ctx->L[0]);
  a.ylive = ctx->d_ylive;
  for (int kb0 = 0; kb0 < ctx->nbin; kb0 += ctx->xwin_sim) {
    a.kb0 = kb0;
    a.nbw = std::min(ctx->xwin_sim, ctx->nbin - kb0);
    HIPCHK(ctx, psh_fft_x_bin(N, a, ctx->tw, ctx->stream));
  }
  ev_record(ctx, E_XDONE);
  HIPCHK(ctx, psh_reduce_partials(ctx->partials, nmock, grid, nacc, out_bins, ctx->stream));
  return PSH_OK;
}

// Normalise the raw sums of a simulation box (src/multipole.c:1073-1083).
void normalise_sim(psh_ctx *ctx, const double *raw, double shot, double norm, double *pl) {
  const int nb = ctx->nbin, nl = ctx->conf.npole;
  for (int l = 0; l < nl; l++)
    for (int i = 0; i < nb; i++) {
      double v = raw[l * nb + i];
      if (ctx->cnt[i]) {
        v /= norm * ctx->cnt[i];
        v -= shot * ctx->lcnt[i + l * nb] / ctx->cnt[i];
      }
      v *= 2 * ctx->conf.poles[l] + 1;
      pl[l * nb + i] = v;
    }
}

void fill_pk_common(psh_ctx *ctx, psh_pk *out) {
  const int nb = ctx->nbin, nl = ctx->conf.npole;
  out->nbin = nb;
  out->nl = nl;
  for (int l = 0; l <= PSH_MAX_ELL; l++) out->poles[l] = l < nl ? ctx->conf.poles[l] : 0;
  if (out->k) memcpy(out->k, ctx->kcen.data(), sizeof(double) * nb);
  if (out->kedge) memcpy(out->kedge, ctx->kedge.data(), sizeof(double) * (nb + 1));
  if (out->km) memcpy(out->km, ctx->km.data(), sizeof(double) * nb);
  if (out->cnt)
    for (int i = 0; i < nb; i++) out->cnt[i] = ctx->cnt[i];
  if (out->lcnt) memcpy(out->lcnt, ctx->lcnt.data(), sizeof(double) * nb * nl);
  // log-bin rescale, as the reference writes it (src/multipole.c:1267-1273):
  // km and the lower edges are exponentiated from the already-rescaled k
  if (ctx->conf.logscale) {
    for (int i = 0; i < nb; i++) {
      if (out->k) out->k[i] = pow(10, out->k[i]);
      double k = out->k ? out->k[i] : pow(10, ctx->kcen[i]);
      if (out->km) out->km[i] = pow(10, k);
      if (out->kedge) out->kedge[i] = pow(10, k);
    }
    if (out->kedge) out->kedge[nb] = pow(10, out->kedge[nb]);
  }
}

// Survey path for catalogue `c`: data and randoms on the device.
// Number of raw bin slots of a survey run: l = 0, then 2l+1 per further pole.
int survey_slots(const psh_ctx *ctx) {
  int nslots = 1;
  for (int n = 1; n < ctx->conf.npole; n++) nslots += 2 * ctx->conf.poles[n] + 1;
  return nslots;
}

// Raw survey sums of the field of `hp` (data parts with weight 1, random
// parts with -alpha; several catalogues' parts give the field of their sum):
// raw[0..nb) = sum |F0|^2 alias mult (src/multipole.c:579-660), then per pole
// n >= 1 and m = -l..l: sum Re(F0 conj(Y_lm(k) S_lm)) with S_lm the transform
// of Y_lm(r) times the field (src/mp_template.c:48-149).
int survey_raw(psh_ctx *ctx, const std::vector<PshDevPart> &hp, std::vector<double> &raw,
               double dc) {
  const Mark mark("psh:survey");
  const int N = ctx->Ng;
  const bool il = ctx->conf.intlace;
  const int nb = ctx->nbin, nl = ctx->conf.npole;
  ev_record(ctx, E_START);
  int rc;
  // plain grid -> A when interlaced (combined into A), else -> B (real field
  // kept for the multipole weights) and transformed out of place into A
  float *plain = il ? ctx->meshA : ctx->meshB;
  if ((rc = run_assign(ctx, hp, 1, plain, il ? ctx->meshB : nullptr))) return rc;

  float2 *A = reinterpret_cast<float2 *>(ctx->meshA);
  float2 *B = reinterpret_cast<float2 *>(ctx->meshB);
  PshYlmR noy{};
  if (il) {
    HIPCHK(ctx, psh_fft_z_r2c(N, ctx->meshA, A, 1, ctx->mstride, ctx->tw, noy, ctx->stream));
    HIPCHK(ctx, psh_fft_z_r2c(N, ctx->meshB, B, 1, ctx->mstride, ctx->tw, noy, ctx->stream));
    HIPCHK(ctx, psh_fft_y(N, A, 1, ctx->mstride, ctx->tw, 0, ctx->stream, nullptr,
                          ctx->ysched));
    HIPCHK(ctx, psh_fft_y(N, B, 1, ctx->mstride, ctx->tw, 0, ctx->stream, nullptr,
                          ctx->ysched));
  } else {
    HIPCHK(ctx, psh_fft_z_r2c(N, ctx->meshB, A, 1, ctx->mstride, ctx->tw, noy, ctx->stream));
    HIPCHK(ctx, psh_fft_y(N, A, 1, ctx->mstride, ctx->tw, 0, ctx->stream, nullptr,
                          ctx->ysched));
  }
  ev_record(ctx, E_YDONE);
  // the l = 0 and l > 0 passes each at their own grid, over one partials buffer
  const int grid = xgrid_for(ctx, 1, ctx->xres_s0);
  const int gridl = xgrid_for(ctx, 1, ctx->xres_sl);
  const int nslots = survey_slots(ctx);
  if ((rc = ensure_partials(ctx, 1, std::max(grid, gridl), nb, nslots))) return rc;
  PshXBinArgs a{};
  a.epi = PSH_EPI_SURVEY0;
  a.A = A;
  a.B = il ? B : nullptr;
  a.store = A;
  a.nmock = 1;
  a.mstride_floats = ctx->mstride;
  a.ax = ctx->ax;
  a.bins = ctx->bins;
  a.phase = ctx->phase;
  a.partials = ctx->partials;
  a.grid = grid;
  a.vec0f = (float)(2 * M_PI / ctx->L[0]);
  a.has_dc = 1;
  a.dc = dc;
  // F0 is stored over its own input (A): the launch that stores it must be
  // the last to read A, so the windows run from the last to the first and
  // only the first window's launch (kb0 = 0) stores
  const int nwin = (nb + ctx->xwin_survey - 1) / ctx->xwin_survey;
  for (int wi = nwin - 1; wi >= 0; wi--) {
    PshXBinArgs w = a;
    w.kb0 = wi * ctx->xwin_survey;
    w.nbw = std::min(ctx->xwin_survey, nb - w.kb0);
    if (wi > 0) w.store = nullptr;
    HIPCHK(ctx, psh_fft_x_bin(N, w, ctx->tw, ctx->stream));
  }
  ev_record(ctx, E_XDONE);
  HIPCHK(ctx, psh_reduce_partials(ctx->partials, 1, grid, nb, ctx->d_bins, ctx->stream));
  if (il) {
    // back to configuration space: B = c2r(F0) / Ntot (src/multipole.c:489-500)
    HIPCHK(ctx, psh_fft_x(N, A, B, 1, ctx->mstride, ctx->tw, 1, ctx->stream));
    HIPCHK(ctx, psh_fft_y(N, B, 1, ctx->mstride, ctx->tw, 1, ctx->stream));
    const double ntot = (double)N * N * N;
    HIPCHK(ctx, psh_fft_z_c2r(N, B, 1, ctx->mstride, ctx->tw, (float)(1.0 / ntot),
                              ctx->stream));
  }
  // multipoles l > 0 (src/multipole.c:1237-1245, src/mp_template.c:48-149);
  // the first pole is skipped when it is not 0, as in the reference
  int slot = 1;
  float2 *S = reinterpret_cast<float2 *>(ctx->meshS);
  const bool lprof = ctx->profiling && !ctx->arena;
  if (lprof)
    while (ctx->tm.lev.size() < (size_t)3 * (nslots - 1)) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) break;
      ctx->tm.lev.push_back(e);
    }
  auto lrec = [&](int q) {
    if (lprof && (size_t)q < ctx->tm.lev.size())
      (void)hipEventRecord(ctx->tm.lev[q], ctx->stream);
  };
  // The (l, m) fields of a pole go in groups of up to ctx->nfs (the S
  // buffers): each field's z and y passes, then one fused x pass that
  // transforms every field of the group and bins Re(F0 conj(sum_m Y_lm(k)
  // F_lm)) once.  A group's sums go to the slot of its first m; the other
  // slots of the pole stay zero (the host adds a pole's slots).
  if (nslots > 1)
    HIPCHK(ctx, hipMemsetAsync(ctx->d_bins + nb, 0, sizeof(double) * (size_t)(nslots - 1) * nb,
                               ctx->stream));
  int ngrp = 0;
  const size_t fstride_c = ctx->mstride / 2;
  for (int n = 1; n < nl; n++) {
    const int ell = ctx->conf.poles[n];
    for (int m0 = -ell; m0 <= ell; m0 += ctx->nfs) {
      const int nf = std::min(ctx->nfs, ell - m0 + 1);
      const int gslot = slot;
      const Mark lm("psh:survey_lm");
      lrec(3 * ngrp);
      PshXBinArgs b = a;
      b.epi = PSH_EPI_SURVEYL;
      b.grid = gridl;
      b.has_dc = 0;
      b.ylive = ctx->d_ylive;   // the S fields are read by this pass only
      b.A = S;
      b.B = nullptr;
      b.store = nullptr;
      b.F0 = A;
      b.nf = nf;
      b.fstride_floats = ctx->mstride;
      for (int f = 0; f < nf; f++, slot++) {
        PshYlmR y{};
        y.ell = ell;
        y.m = m0 + f;
        y.norm = ylm_norm(ell, y.m);
        // r-space position of cell i: (i + smin_g) L with smin_g = bmin Ng / L
        // (src/genr_mesh.c:913-914, src/mp_template.c:71-83), i L + bmin Ng
        for (int ax = 0; ax < 3; ax++) {
          y.org[ax] = (float)(ctx->bmin[ax] * N);
          y.step[ax] = (float)ctx->L[ax];
        }
        float2 *Sf = S + (size_t)f * fstride_c;
        HIPCHK(ctx, psh_fft_z_r2c(N, ctx->meshB, Sf, 1, ctx->mstride, ctx->tw, y, ctx->stream,
                                  ctx->kzlim));
        HIPCHK(ctx, psh_fft_y(N, Sf, 1, ctx->mstride, ctx->tw, 0, ctx->stream, ctx->d_ylive,
                              ctx->ysched));
        b.ells[f] = ell;
        b.ms[f] = y.m;
        b.ynorms[f] = y.norm;
      }
      lrec(3 * ngrp + 1);
      for (int kb0 = 0; kb0 < nb; kb0 += ctx->xwin_survey) {
        b.kb0 = kb0;
        b.nbw = std::min(ctx->xwin_survey, nb - kb0);
        HIPCHK(ctx, psh_fft_x_bin(N, b, ctx->tw, ctx->stream));
      }
      lrec(3 * ngrp + 2);
      ngrp++;
      HIPCHK(ctx, psh_reduce_partials(ctx->partials, 1, gridl, nb,
                                      ctx->d_bins + (size_t)gslot * nb, ctx->stream));
    }
  }
  ev_record(ctx, E_END);
  raw.assign((size_t)nslots * nb, 0.0);
  HIPCHK(ctx, hipMemcpyAsync(raw.data(), ctx->d_bins, sizeof(double) * raw.size(),
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ev_collect(ctx);
  if (ctx->profiling) {
    // surveys: [5] the l > 0 fields' fused x passes, [6] their z and y passes
    double tx = 0.0, tzy = 0.0;
    const size_t nev = ctx->tm.lev.size();
    bool ok = lprof && nev >= (size_t)3 * ngrp;
    for (int q = 0; ok && q < ngrp; q++) {
      float a0 = 0.f, a1 = 0.f;
      ok = hipEventElapsedTime(&a0, ctx->tm.lev[3 * q], ctx->tm.lev[3 * q + 1]) == hipSuccess &&
           hipEventElapsedTime(&a1, ctx->tm.lev[3 * q + 1], ctx->tm.lev[3 * q + 2]) == hipSuccess;
      if (!ok) {
        (void)hipGetLastError();
        break;
      }
      tzy += a0;
      tx += a1;
    }
    ctx->ms[5] = ok ? tx : -1.0;
    ctx->ms[6] = ok ? tzy : -1.0;
  }
  return PSH_OK;
}

// Survey normalisation (src/multipole.c:1096-1153): l = 0 as (S/cnt - shot)
// / norm when poles[0] == 0, l > 0 as S * 4pi / (norm cnt) summed over m.
// Cross spectra (shot = 0, norm = sqrt(norm0 norm1), fac = 2pi for l > 0 over
// the two orderings) reuse it with cross = true.
void normalise_survey(const psh_ctx *ctx, const double *raw, double shot, double norm,
                      bool cross, double *pl) {
  const int nb = ctx->nbin, nl = ctx->conf.npole;
  for (int q = 0; q < nl * nb; q++) pl[q] = 0;
  if (ctx->conf.poles[0] == 0) {   // src/multipole.c:1110-1114
    for (int i = 0; i < nb; i++) {
      double v = raw[i];
      if (ctx->cnt[i]) {
        if (cross) {
          v /= norm * ctx->cnt[i];
        } else {
          v = v / ctx->cnt[i] - shot;
          v /= norm;
        }
      }
      pl[i] = v;
    }
  }
  int slot = 1;
  for (int n = 1; n < nl; n++) {   // src/multipole.c:1146-1153
    const int ell = ctx->conf.poles[n];
    std::vector<double> s(nb, 0.0);
    for (int m = -ell; m <= ell; m++, slot++)
      for (int i = 0; i < nb; i++) s[i] += raw[(size_t)slot * nb + i];
    for (int i = 0; i < nb; i++) {
      double v = s[i];
      if (ctx->cnt[i]) {
        double nrm = norm * ctx->cnt[i];
        v *= (cross ? 2 : 4) * M_PI / nrm;
      }
      pl[n * nb + i] = v;
    }
  }
}

// Coordinate bounds of n device records folded into mn/mx; *wsum (may be
// NULL) receives the sum of their weights (block sums added in block order).
int bounds_device(psh_ctx *ctx, const double *d, size_t n, double mn[3], double mx[3],
                  double *wsum = nullptr) {
  if (wsum) *wsum = 0.0;
  if (n == 0) return PSH_OK;
  int blocks = (int)((n + 255) / 256);
  if (blocks > 1024) blocks = 1024;
  int rc;
  if (!ctx->d_bounds &&
      (rc = dev_alloc(ctx, &ctx->d_bounds, (size_t)1024 * PSH_BOUNDS_STRIDE)))
    return rc;
  HIPCHK(ctx, psh_coord_bounds(d, n, ctx->d_bounds, blocks, ctx->stream));
  std::vector<double> h((size_t)blocks * PSH_BOUNDS_STRIDE);
  HIPCHK(ctx, hipMemcpyAsync(h.data(), ctx->d_bounds, sizeof(double) * h.size(),
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  double ws = 0.0;
  for (int b = 0; b < blocks; b++) {
    for (int a = 0; a < 3; a++) {
      if (mn[a] > h[b * PSH_BOUNDS_STRIDE + a]) mn[a] = h[b * PSH_BOUNDS_STRIDE + a];
      if (mx[a] < h[b * PSH_BOUNDS_STRIDE + 3 + a]) mx[a] = h[b * PSH_BOUNDS_STRIDE + 3 + a];
    }
    ws += h[b * PSH_BOUNDS_STRIDE + 6];
  }
  if (wsum) *wsum = ws;
  return PSH_OK;
}

// Sum of the weights of n host records, in record order.
// One pass over host records: coordinate bounds and the sum of the
// weights.  The records are cut into kScanPieces fixed pieces scanned by up
// to 16 host threads (one 0.64 GB catalogue: about 10 ms instead of the
// 100 ms of two single-threaded passes); the pieces' sums are added in piece
// order, so the result does not depend on the thread count.
struct HostScan {
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
  double mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  double wsum = 0.0;
};

constexpr int kScanPieces = 64;

HostScan scan_host(const psh_data *d, size_t n) {
  HostScan piece[kScanPieces];
  auto run = [&](int q) {
    const size_t lo = n * q / kScanPieces, hi = n * (q + 1) / kScanPieces;
    HostScan &s = piece[q];
    for (size_t p = lo; p < hi; p++) {
      for (int a = 0; a < 3; a++) {
        const double v = d[p].x[a];
        if (s.mn[a] > v) s.mn[a] = v;
        if (s.mx[a] < v) s.mx[a] = v;
      }
      s.wsum += d[p].w;
    }
  };
  const unsigned hw = std::thread::hardware_concurrency();
  const int nt = n < ((size_t)1 << 20) ? 1 : (int)std::min(16u, std::max(1u, hw));
  if (nt == 1) {
    for (int q = 0; q < kScanPieces; q++) run(q);
  } else {
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; t++)
      th.emplace_back([&, t]() {
        for (int q = t; q < kScanPieces; q += nt) run(q);
      });
    for (auto &x : th) x.join();
  }
  HostScan r;
  for (int q = 0; q < kScanPieces; q++) {
    for (int a = 0; a < 3; a++) {
      if (r.mn[a] > piece[q].mn[a]) r.mn[a] = piece[q].mn[a];
      if (r.mx[a] < piece[q].mx[a]) r.mx[a] = piece[q].mx[a];
    }
    r.wsum += piece[q].wsum;
  }
  return r;
}

// Bounds of a catalogue set (data, and randoms of surveys) from the scans of
// its arrays; sd[c] / sr[c] keep each array's weight sum.
void scan_cat(const psh_ctx *ctx, const psh_cata *cat, double mn[3], double mx[3],
              double sd[2], double sr[2]) {
  for (int a = 0; a < 3; a++) {
    mn[a] = DBL_MAX;
    mx[a] = -DBL_MAX;
  }
  auto merge = [&](const HostScan &s) {
    for (int a = 0; a < 3; a++) {
      if (mn[a] > s.mn[a]) mn[a] = s.mn[a];
      if (mx[a] < s.mx[a]) mx[a] = s.mx[a];
    }
  };
  for (int c = 0; c < cat->num && c < 2; c++) {
    const HostScan s = scan_host(cat->data[c], cat->ndata[c]);
    merge(s);
    sd[c] = s.wsum;
    sr[c] = 0.0;
    if (!ctx->conf.issim) {
      const HostScan r = scan_host(cat->rand[c], cat->nrand[c]);
      merge(r);
      sr[c] = r.wsum;
    }
  }
}

// Host records to device memory through two pinned chunks, allocated on the
// first upload of the process and kept (pinned host memory serves every
// device).  Host threads fill one chunk while the other's DMA runs on the
// context's stream; the call returns when the copy is complete, so the
// caller may free the source.  The runtime's pageable path ran at 7 to 27 GB/s
// on MI355X depending on the source buffer's history (a fresh mapping at an
// address an earlier, freed catalogue had used was the slow case).
constexpr size_t kUpChunk = (size_t)32 << 20;
std::mutex g_up_mu;
void *g_up_buf[2] = {nullptr, nullptr};

int upload_host(psh_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return PSH_OK;
  std::lock_guard<std::mutex> lk(g_up_mu);
  for (int b = 0; b < 2; b++)
    if (!g_up_buf[b] &&
        hipHostMalloc(&g_up_buf[b], kUpChunk, hipHostMallocPortable) != hipSuccess) {
      g_up_buf[b] = nullptr;
      (void)hipGetLastError();
      // no pinned staging: the runtime's own pageable copy
      HIPCHK(ctx, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
      return PSH_OK;
    }
  hipEvent_t ev[2] = {nullptr, nullptr};
  int rc = PSH_OK;
  for (int b = 0; b < 2 && !rc; b++)
    if (hipEventCreateWithFlags(&ev[b], hipEventDisableTiming) != hipSuccess)
      rc = set_err(ctx, PSH_ERR_DEVICE, "event creation failed");
  const unsigned hw = std::thread::hardware_concurrency();
  const int nt = (int)std::min(8u, std::max(1u, hw));
  bool used[2] = {false, false};
  for (size_t off = 0, i = 0; off < bytes && !rc; off += kUpChunk, i++) {
    const int b = (int)(i & 1);
    const size_t len = std::min(kUpChunk, bytes - off);
    // the chunk's previous DMA must be done before it is refilled
    if (used[b] && hipEventSynchronize(ev[b]) != hipSuccess) {
      rc = set_err(ctx, PSH_ERR_DEVICE, "upload wait failed");
      break;
    }
    char *buf = static_cast<char *>(g_up_buf[b]);
    const char *s = static_cast<const char *>(src) + off;
    if (nt == 1 || len < ((size_t)1 << 20)) {
      memcpy(buf, s, len);
    } else {
      std::vector<std::thread> th;
      th.reserve(nt);
      for (int t = 0; t < nt; t++)
        th.emplace_back([=]() {
          const size_t lo = len * t / nt, hi = len * (t + 1) / nt;
          memcpy(buf + lo, s + lo, hi - lo);
        });
      for (auto &x : th) x.join();
    }
    if (hipMemcpyAsync(static_cast<char *>(dst) + off, buf, len, hipMemcpyHostToDevice,
                       ctx->stream) != hipSuccess ||
        hipEventRecord(ev[b], ctx->stream) != hipSuccess) {
      rc = set_err(ctx, PSH_ERR_DEVICE, "upload copy failed");
      break;
    }
    used[b] = true;
  }
  // the staging chunks are shared: every DMA from them ends before the lock
  // is released, also after an error
  if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc)
    rc = set_err(ctx, PSH_ERR_DEVICE, "upload failed");
  for (int b = 0; b < 2; b++)
    if (ev[b]) (void)hipEventDestroy(ev[b]);
  return rc;
}

int stage_host(psh_ctx *ctx, const psh_data *h, size_t n, size_t offset_records) {
  if (n == 0) return PSH_OK;
  return upload_host(ctx, ctx->d_cat + 4 * offset_records, h, sizeof(psh_data) * n);
}

}  // namespace

// ===========================================================================
// context layer
// ===========================================================================

namespace {

// Set up a context for conf on the current device.  With ctx->arena set the
// buffers come from the arena and the work runs on the null stream (no
// stream or event is created); otherwise on a stream of its own.  On error
// ctx->err holds the message and the caller destroys the context.
int ctx_init(psh_ctx *ctx, const psh_conf *conf, int nmock_max) {
  int rc = validate_conf(conf, ctx);
  if (rc) return rc;
  ctx->conf = *conf;
  ctx->nmock_max = nmock_max < 1 ? 1 : nmock_max;
  ctx->Ng = conf->gsize;
  ctx->M = ctx->Ng / 2;
  ctx->H = ctx->M + PSH_PAD;
  ctx->mstride = (size_t)ctx->Ng * ctx->Ng * 2 * ctx->H;
  ctx->survey = !conf->issim;
  ctx->ellmax = conf->poles[conf->npole - 1];
  if (ctx->survey && ctx->nmock_max > 1)
    return set_err(ctx, PSH_ERR_UNSUPPORTED, "batched mocks are supported for simulation boxes");
  if (hipGetDevice(&ctx->dev) != hipSuccess ||
      (!ctx->arena &&
       hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)) {
    ctx->stream = nullptr;
    return set_err(ctx, PSH_ERR_DEVICE, "no usable HIP device");
  }
  const size_t nm = (size_t)ctx->nmock_max;
  if ((rc = dev_alloc(ctx, &ctx->meshA, ctx->mstride * nm)) ||
      ((conf->intlace || ctx->survey) && (rc = dev_alloc(ctx, &ctx->meshB, ctx->mstride * nm))) ||
      (ctx->survey && ctx->ellmax > 0 && (rc = alloc_survey_fields(ctx))))
    return rc;
  if ((rc = dev_alloc(ctx, &ctx->tw, (size_t)ctx->Ng))) return rc;
  if (psh_fft_twiddles_fill(ctx->Ng, ctx->tw) != hipSuccess)
    return set_err(ctx, PSH_ERR_DEVICE, "twiddle table upload failed");
  {
    std::vector<float2> ph(ctx->Ng);
    for (int i = 0; i < ctx->Ng; i++) {
      int s = (i <= ctx->M) ? i : i - ctx->Ng;
      double a = M_PI * s / ctx->Ng;
      ph[i] = make_float2((float)cos(a), (float)sin(a));
    }
    if ((rc = dev_alloc(ctx, &ctx->phase, (size_t)ctx->Ng))) return rc;
    if (hipMemcpy(ctx->phase, ph.data(), sizeof(float2) * ctx->Ng, hipMemcpyHostToDevice) !=
        hipSuccess)
      return set_err(ctx, PSH_ERR_DEVICE, "phase table upload failed");
  }
  if (!ctx->arena)
    for (int i = 0; i < E_COUNT; i++) (void)hipEventCreate(&ctx->tm.ev[i]);
  if (conf->issim) {
    double bmin[3] = {0, 0, 0};
    if ((rc = set_box(ctx, conf->bsize, bmin))) return rc;
  }
  return PSH_OK;
}

}  // namespace

extern "C" int psh_ctx_create(const psh_conf *conf, int nmock_max, psh_ctx **out) {
  if (!out) return set_err(nullptr, PSH_ERR_ARG, "null output pointer");
  *out = nullptr;
  psh_ctx *ctx = new (std::nothrow) psh_ctx();
  if (!ctx) return set_err(nullptr, PSH_ERR_MEMORY, "out of host memory");
  int rc = ctx_init(ctx, conf, nmock_max);
  if (rc) {
    const std::string e = ctx->err;
    psh_ctx_destroy(ctx);
    return set_err(nullptr, rc, "%s", e.c_str());
  }
  *out = ctx;
  return PSH_OK;
}

extern "C" int psh_ctx_destroy(psh_ctx *ctx) {
  if (!ctx) return PSH_OK;
  if (ctx->stream || ctx->arena) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  void *bufs[] = {ctx->meshA, ctx->meshB, ctx->meshS, ctx->counts, ctx->starts, ctx->cursors,
                  ctx->scan_tmp, ctx->d_parts, ctx->records, ctx->d_flags, ctx->tw,
                  ctx->phase, ctx->d_tab, ctx->d_ylive, ctx->d_gcnt, ctx->d_gkm,
                  ctx->d_glcnt, ctx->partials, ctx->d_bins, ctx->d_bounds, ctx->d_cat,
                  ctx->whist, ctx->d_mlp, ctx->d_ax, ctx->d_ay, ctx->d_acorr, ctx->d_afeat};
  for (void *p : bufs) ctx_free(ctx, p);
  for (int i = 0; i < E_COUNT; i++)
    if (ctx->tm.ev[i]) (void)hipEventDestroy(ctx->tm.ev[i]);
  for (hipEvent_t e : ctx->tm.lev) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return PSH_OK;
}

extern "C" int psh_ctx_set_box(psh_ctx *ctx, const double bsize[3], const double bmin[3]) {
  if (!ctx || !bsize || !bmin) return set_err(ctx, PSH_ERR_ARG, "null argument");
  for (int a = 0; a < 3; a++) ctx->cmax[a] = bmin[a] + bsize[a];
  return set_box(ctx, bsize, bmin);
}

extern "C" int psh_ctx_nbin(const psh_ctx *ctx) { return ctx && ctx->box_set ? ctx->nbin : 0; }

extern "C" int psh_ctx_set_profiling(psh_ctx *ctx, int on) {
  if (!ctx) return PSH_ERR_ARG;
  ctx->profiling = on != 0;
  return PSH_OK;
}

extern "C" int psh_ctx_timings(const psh_ctx *ctx, double ms[8]) {
  if (!ctx || !ms) return PSH_ERR_ARG;
  for (int i = 0; i < 8; i++) ms[i] = ctx->ms[i];
  return PSH_OK;
}

extern "C" const char *psh_last_error(const psh_ctx *ctx) {
  if (ctx) return ctx->err.c_str();
  return g_last_error.c_str();
}

extern "C" int psh_ctx_set_mu_cells(psh_ctx *ctx, int nmu) {
  if (!ctx) return set_err(nullptr, PSH_ERR_ARG, "null context");
  if (ctx->survey) return set_err(ctx, PSH_ERR_ARG, "|mu| cells are for simulation boxes");
  if (nmu < 1 || nmu > 64) return set_err(ctx, PSH_ERR_ARG, "nmu %d outside [1, 64]", nmu);
  if (nmu == ctx->nmu) return PSH_OK;
  ctx->nmu = nmu;
  ctx->cells.clear();
  if (ctx->box_set) {
    // the bins' accumulator layout changes: redo the binning set-up (the
    // tables and geometry stay); a captured graph holds the old layout
    drop_graph(ctx);
    int rc = make_bins(ctx);
    if (rc) {
      ctx->box_set = false;
      ctx->geom_valid = false;
      return rc;
    }
  }
  return PSH_OK;
}

extern "C" int psh_ctx_cells(const psh_ctx *ctx, double *cells) {
  if (!ctx || !cells) return set_err(nullptr, PSH_ERR_ARG, "null argument");
  if (ctx->cells.empty())
    return set_err(const_cast<psh_ctx *>(ctx), PSH_ERR_ARG, "no box run with cells yet");
  memcpy(cells, ctx->cells.data(), sizeof(double) * ctx->cells.size());
  return PSH_OK;
}

namespace {

int ensure_alias_rows(psh_ctx *ctx) {
  const size_t rows = (size_t)ctx->nbin * ctx->bins.nmu;
  int rc;
  if (rows > ctx->arows_cap) {
    ctx->arows_cap = 0;
    if ((rc = dev_alloc(ctx, &ctx->d_ax, rows * 32))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_ay, rows * (size_t)std::max(ctx->mlp_nnet, 1)))) return rc;
    ctx->arows_cap = rows;
  }
  // per-bin buffers follow the bin count, whatever the number of cells
  if (ctx->nbin > ctx->afeat_nbin) {
    ctx->afeat_nbin = 0;
    if ((rc = dev_alloc(ctx, &ctx->d_acorr, (size_t)ctx->conf.npole * ctx->nbin))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_afeat, psh_alias_scratch_doubles(ctx->nbin)))) return rc;
    ctx->afeat_nbin = ctx->nbin;
  }
  return PSH_OK;
}

}  // namespace

extern "C" int psh_ctx_alias_rows(psh_ctx *ctx, float *X) {
  if (!ctx || !X) return set_err(ctx, PSH_ERR_ARG, "null argument");
  if (ctx->cells.empty() || ctx->bins.nmu < 1 || ctx->conf.poles[0] != 0)
    return set_err(ctx, PSH_ERR_ARG, "no box run with |mu| cells and a monopole yet");
  int rc;
  if ((rc = ensure_alias_rows(ctx))) return rc;
  const int nb = ctx->nbin, nmu = ctx->bins.nmu;
  HIPCHK(ctx, hipMemcpyAsync(ctx->d_bins, ctx->cells.data(), sizeof(double) * nb * nmu,
                             hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, psh_alias_features(ctx->d_bins, ctx->d_gcnt, nb, nmu, ctx->last_norm,
                                 ctx->last_shot,
                                 M_PI * ctx->Ng / std::max(ctx->L[0], std::max(ctx->L[1], ctx->L[2])),
                                 ctx->bins.kedge0, ctx->bins.dk, ctx->bins.logscale,
                                 ctx->conf.assign, ctx->d_ax, ctx->d_afeat, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(X, ctx->d_ax, sizeof(float) * 32 * nb * nmu,
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return PSH_OK;
}

extern "C" int psh_ctx_set_alias_mlp(psh_ctx *ctx, const float *params, int nnet) {
  if (!ctx) return set_err(nullptr, PSH_ERR_ARG, "null context");
  if (!params || nnet <= 0) {
    ctx->mlp_nnet = 0;
    return PSH_OK;
  }
  if (ctx->survey || ctx->conf.intlace || ctx->conf.poles[0] != 0 || ctx->nmu < 2)
    return set_err(ctx, PSH_ERR_ARG,
                   "the learned alias correction is for plain (non-interlaced) boxes with a "
                   "monopole and |mu| cells (psh_ctx_set_mu_cells)");
  if (nnet > 64) return set_err(ctx, PSH_ERR_ARG, "at most 64 networks");
  const size_t P = (size_t)psh_mlp_param_count();
  int rc;
  ctx->mlp_nnet = 0;
  ctx->arows_cap = 0;   // d_ay holds one output row block per network
  if ((rc = dev_alloc(ctx, &ctx->d_mlp, P * nnet))) return rc;
  HIPCHK(ctx, hipMemcpy(ctx->d_mlp, params, sizeof(float) * P * nnet, hipMemcpyHostToDevice));
  ctx->mlp_nnet = nnet;
  if (ctx->box_set && (rc = ensure_alias_rows(ctx))) return rc;
  return PSH_OK;
}

extern "C" int psh_fault_inject(int what, long value) {
  fault_env_init();
  if (what == PSH_FAULT_ALLOC) {
    g_fault_alloc = value;
    return PSH_OK;
  }
  if (what == PSH_FAULT_GRAPH) {
    g_fault_graph = value != 0;
    return PSH_OK;
  }
  return set_err(nullptr, PSH_ERR_ARG, "unknown fault kind %d", what);
}

// Raw box sums of the field of `hp` (several parts: the field of their sum):
// sum |F|^2 alias mult L_l(mu) per pole and bin (src/multipole.c:778-898),
// after the box check of every particle (src/genr_mesh.c:517-533).
// The box of a simulation comes from the configuration, so the check of
// every coordinate against it rides on the count pass (flag bit 4); only a
// failed check runs the separate min/max pass, for the reference's message.
int sim_box_error(psh_ctx *ctx, const std::vector<PshDevPart> &hp) {
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  int rc;
  for (auto &p : hp)
    if ((rc = bounds_device(ctx, p.xyzw, p.n, mn, mx))) return rc;
  if ((rc = def_box(ctx, mn, mx))) return rc;
  return set_err(ctx, PSH_ERR_CATA, "particle coordinate outside the box [0, BOX_SIZE)");
}

// Raw box sums of the field of `hp` per pole and bin (raw[nl][nbin]); the
// per-cell sums (|mu| cells) stay in ctx->cells.  With a learned alias
// correction set (plain grids) and `correct`, the sums are corrected on the
// device: feature rows from the run's own cell sums (norm and shot of the
// catalogue), the networks' outputs, and sum_c exp(f) S[l][b][c].
int sim_raw(psh_ctx *ctx, const std::vector<PshDevPart> &hp, std::vector<double> &raw,
            double norm = 0, double shot = 0, bool correct = false) {
  int rc;
  ev_record(ctx, E_START);
  size_t ntot = 0;
  for (auto &p : hp) ntot += p.n;
  if (ntot == 0) return sim_box_error(ctx, hp);   // "invalid coordinate" as the reference
  const double zero[3] = {0, 0, 0};
  if ((rc = def_box(ctx, zero, zero))) return rc;
  if ((rc = ensure_geometry(ctx))) return rc;
  rc = run_assign(ctx, hp, 1, ctx->meshA, ctx->conf.intlace ? ctx->meshB : nullptr);
  if (rc == PSH_ERR_CATA) return sim_box_error(ctx, hp);
  if (rc) return rc;
  const int nacc = psh_xbin_nacc(ctx->bins, PSH_EPI_SIM);
  if ((rc = ensure_partials(ctx, 1, xgrid_for(ctx, 1, ctx->xres_sim), nacc, 1))) return rc;
  if ((rc = run_sim_fft_bin(ctx, 1, ctx->d_bins))) return rc;
  const int nb = ctx->nbin, nl = ctx->conf.npole, nmu = ctx->bins.nmu;
  const bool mlp = correct && ctx->mlp_nnet > 0 && !ctx->conf.intlace && nmu > 1;
  ctx->last_norm = norm;
  ctx->last_shot = shot;
  if (mlp && (rc = ensure_alias_rows(ctx))) return rc;
  if (mlp) {
    const int rows = nb * nmu;
    HIPCHK(ctx, psh_alias_features(ctx->d_bins, ctx->d_gcnt, nb, nmu, norm, shot,
                                   M_PI * ctx->Ng / std::max(ctx->L[0], std::max(ctx->L[1], ctx->L[2])),
                                   ctx->bins.kedge0, ctx->bins.dk, ctx->bins.logscale,
                                   ctx->conf.assign, ctx->d_ax, ctx->d_afeat, ctx->stream));
    // the ensemble in one launch (the networks' rows side by side in d_ay)
    HIPCHK(ctx, psh_mlp_forward_async(ctx->d_mlp, ctx->d_ax, rows, ctx->d_ay, ctx->stream,
                                      ctx->mlp_nnet));
    HIPCHK(ctx, psh_alias_apply(ctx->d_ay, ctx->mlp_nnet, ctx->d_ax, ctx->d_bins, nl, nb, nmu,
                                ctx->d_acorr, ctx->stream));
  }
  ev_record(ctx, E_END);
  ctx->cells.assign(nacc, 0.0);
  HIPCHK(ctx, hipMemcpyAsync(ctx->cells.data(), ctx->d_bins, sizeof(double) * nacc,
                             hipMemcpyDeviceToHost, ctx->stream));
  raw.assign((size_t)nl * nb, 0.0);
  if (mlp)
    HIPCHK(ctx, hipMemcpyAsync(raw.data(), ctx->d_acorr, sizeof(double) * nl * nb,
                               hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (!mlp)
    for (size_t q = 0; q < raw.size(); q++)
      for (int c = 0; c < nmu; c++) raw[q] += ctx->cells[q * nmu + c];
  ev_collect(ctx);
  return PSH_OK;
}

// Cross normalisation of boxes (src/multipole.c:1048-1061): no shot noise,
// divided by sqrt(norm0 norm1) cnt, times 2l+1.
void normalise_sim_cross(const psh_ctx *ctx, const double *raw, double norm, double *pl) {
  const int nb = ctx->nbin, nl = ctx->conf.npole;
  for (int l = 0; l < nl; l++)
    for (int i = 0; i < nb; i++) {
      double v = raw[l * nb + i];
      if (ctx->cnt[i]) v /= norm * ctx->cnt[i];
      pl[l * nb + i] = v * (2 * ctx->conf.poles[l] + 1);
    }
}

extern "C" int psh_ctx_run_device(psh_ctx *ctx, const psh_data *d_data, size_t n,
                                  double weight_sum, psh_pk *out) {
  if (!ctx || !out) return set_err(ctx, PSH_ERR_ARG, "null argument");
  if (ctx->survey)
    return set_err(ctx, PSH_ERR_ARG, "psh_ctx_run_device is for simulation boxes");
  int rc;
  std::vector<PshDevPart> hp(1);
  hp[0] = PshDevPart{(const double *)d_data, n, 0, 1.0f, 0};
  std::vector<double> raw;
  // shot noise and normalisation of a box (src/genr_mesh.c:904-910)
  const double vol = ctx->L[0] * ctx->L[1] * ctx->L[2];
  const double shot = vol / weight_sum;
  const double norm = weight_sum * weight_sum / vol;
  if ((rc = sim_raw(ctx, hp, raw, norm, shot, true))) return rc;
  fill_pk_common(ctx, out);
  if (out->pl[0]) normalise_sim(ctx, raw.data(), shot, norm, out->pl[0]);
  return PSH_OK;
}

// ===========================================================================
// allocation-free entry point
// ===========================================================================

// Device scratch of psh_run_device_static: one array per GPU, held by the
// process from the library's load to its end.  Host code never passes its
// name to a kernel or a copy (on the host it only names the symbol): the
// device address comes from hipGetSymbolAddress, once per device, below.
static __device__ __align__(256) char scratch[PSH_STATIC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (dev).
int scratch_address(int dev, char **out) {
  if (dev < 0 || dev >= kMaxDevices)
    return set_err(nullptr, PSH_ERR_DEVICE, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(nullptr, PSH_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                     hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<char *>(p);
  }
  *out = g_scratch_ptr[dev];
  return PSH_OK;
}

}  // namespace

extern "C" int psh_run_device_static(const psh_conf *conf, const psh_data *d_data, size_t n,
                                     double weight_sum, psh_pk *out) {
  if (!conf || !out || (n && !d_data)) return set_err(nullptr, PSH_ERR_ARG, "null argument");
  if (!conf->issim)
    return set_err(nullptr, PSH_ERR_ARG, "psh_run_device_static is for simulation boxes");
  if (conf->ndata != 1 || conf->iscross)
    return set_err(nullptr, PSH_ERR_ARG, "psh_run_device_static takes one catalogue");
  int dev = 0, rc;
  if (hipGetDevice(&dev) != hipSuccess) return set_err(nullptr, PSH_ERR_DEVICE, "no HIP device");
  char *base = nullptr;
  if ((rc = scratch_address(dev, &base))) return rc;
  psh_ctx *ctx = new (std::nothrow) psh_ctx();
  if (!ctx) return set_err(nullptr, PSH_ERR_MEMORY, "out of host memory");
  ctx->arena = base;
  ctx->arena_cap = PSH_STATIC_SCRATCH_BYTES;
  const int cap = out->nbin;
  rc = ctx_init(ctx, conf, 1);
  if (!rc && ctx->nbin > cap)
    rc = set_err(ctx, PSH_ERR_ARG, "the output holds %d bins, the run has %d", cap, ctx->nbin);
  if (!rc) rc = psh_ctx_run_device(ctx, d_data, n, weight_sum, out);
  if (rc) {
    const std::string e = ctx->err;
    psh_ctx_destroy(ctx);
    return set_err(nullptr, rc, "%s", e.c_str());
  }
  psh_ctx_destroy(ctx);
  return PSH_OK;
}

// One or two catalogues already on the device: auto spectra of those with
// isauto set, and the cross spectrum when iscross is set (src/multipole.c:
// 1209-1245).  The cross sums come from the field of the two catalogues'
// sum: with S[.] the raw sums of a field, the l = 0 and box terms
// Re(F_0 F_1*) = (S[0+1] - S[0] - S[1]) / 2, and the survey l > 0 terms,
// which the reference adds over both orderings, S[0+1] - S[0] - S[1].
struct CatDev {
  const double *data, *rand;   // device psh_data records
  size_t nd, nr;
  double wdata, alpha, shot, norm;
  double sw_d, sw_r;           // sums of the records' weights (surveys)
};

int run_catalogues(psh_ctx *ctx, const CatDev *cats, int num, psh_pk *out) {
  int rc;
  const bool cross = ctx->conf.iscross && num == 2;
  std::vector<double> raw[2], rawx;
  for (int c = 0; c < num; c++) {
    if (!ctx->conf.isauto[c] && !cross) continue;
    std::vector<PshDevPart> hp;
    hp.push_back(PshDevPart{cats[c].data, cats[c].nd, 0, 1.0f, 0});
    if (ctx->survey) {
      hp.push_back(PshDevPart{cats[c].rand, cats[c].nr, cats[c].nd, (float)(-cats[c].alpha), 0});
      if ((rc = survey_raw(ctx, hp, raw[c], cats[c].sw_d - cats[c].alpha * cats[c].sw_r)))
        return rc;
    } else {
      if ((rc = sim_raw(ctx, hp, raw[c], cats[c].norm, cats[c].shot, !cross))) return rc;
    }
  }
  if (cross) {
    std::vector<PshDevPart> hp;
    size_t start = 0;
    for (int c = 0; c < 2; c++) {
      hp.push_back(PshDevPart{cats[c].data, cats[c].nd, start, 1.0f, 0});
      start += cats[c].nd;
      if (ctx->survey) {
        hp.push_back(PshDevPart{cats[c].rand, cats[c].nr, start, (float)(-cats[c].alpha), 0});
        start += cats[c].nr;
      }
    }
    double dcx = 0.0;
    for (int c = 0; c < 2; c++) dcx += cats[c].sw_d - cats[c].alpha * cats[c].sw_r;
    if ((rc = ctx->survey ? survey_raw(ctx, hp, rawx, dcx) : sim_raw(ctx, hp, rawx))) return rc;
    const size_t nb = ctx->nbin;
    for (size_t q = 0; q < rawx.size(); q++) {
      const double d = rawx[q] - raw[0][q] - raw[1][q];
      // l = 0 slot of a survey and every box term carry Re(F_0 F_1*) once
      rawx[q] = (!ctx->survey || q < nb) ? 0.5 * d : d;
    }
  }
  fill_pk_common(ctx, out);
  for (int c = 0; c < num; c++) {
    if (!ctx->conf.isauto[c] || !out->pl[c]) continue;
    if (ctx->survey)
      normalise_survey(ctx, raw[c].data(), cats[c].shot, cats[c].norm, false, out->pl[c]);
    else
      normalise_sim(ctx, raw[c].data(), cats[c].shot, cats[c].norm, out->pl[c]);
  }
  if (cross && out->xpl) {
    const double nrm = sqrt(cats[0].norm * cats[1].norm);
    if (ctx->survey)
      normalise_survey(ctx, rawx.data(), 0.0, nrm, true, out->xpl);
    else
      normalise_sim_cross(ctx, rawx.data(), nrm, out->xpl);
  }
  return PSH_OK;
}

extern "C" int psh_ctx_run_batch(psh_ctx *ctx, const psh_data *d_data,
                                 const size_t *offsets, const double *weight_sum,
                                 int nmock, int graph, psh_pk *out) {
  const Mark mark("psh:batch");
  if (!ctx || !d_data || !offsets || !weight_sum || !out)
    return set_err(ctx, PSH_ERR_ARG, "null argument");
  if (ctx->survey) return set_err(ctx, PSH_ERR_ARG, "batched runs are for simulation boxes");
  if (nmock < 1 || nmock > ctx->nmock_max)
    return set_err(ctx, PSH_ERR_ARG, "nmock %d outside [1, %d]", nmock, ctx->nmock_max);
  for (int m = 0; m < nmock; m++)
    if (offsets[m + 1] < offsets[m])
      return set_err(ctx, PSH_ERR_ARG, "offsets must be non-decreasing (mock %d)", m);
  if (graph && ctx->arena)
    return set_err(ctx, PSH_ERR_ARG, "no graph capture on the null stream");
  int rc;
  if ((rc = ensure_geometry(ctx))) return rc;
  std::vector<PshDevPart> hp(nmock);
  for (int m = 0; m < nmock; m++)
    hp[m] = PshDevPart{(const double *)d_data + 4 * offsets[m], offsets[m + 1] - offsets[m],
                       offsets[m], 1.0f, m};
  const size_t ntotal = offsets[nmock] - offsets[0];
  for (int m = 0; m < nmock; m++) hp[m].start = offsets[m] - offsets[0];
  const int grid = xgrid_for(ctx, nmock, ctx->xres_sim);
  const int nacc = psh_xbin_nacc(ctx->bins, PSH_EPI_SIM);
  if ((rc = ensure_partials(ctx, nmock, grid, nacc, 1))) return rc;
  if ((rc = ensure_assign_scratch(ctx, nmock, nmock, ntotal))) return rc;
  // boxes: weights of one sign, 64-bit cells only for the fused shape's
  // small grids (run_assign sets the same for the eager path)
  ctx->ag.wide = ctx->ag.zfuse && psh_zfuse_wide(ctx->Ng);
  float *mB = ctx->conf.intlace ? ctx->meshB : nullptr;
  std::vector<double> raw((size_t)nmock * nacc);
  int flags = 0;
  // a replay whose records overflow the buffer is rerun after the buffer
  // grows (the graph, which holds the old buffer, is captured again), as the
  // eager path's run_assign does
  for (int attempt = 0;; attempt++) {
    ev_record(ctx, E_START);
    if (!graph) {
      if ((rc = run_assign(ctx, hp, nmock, ctx->meshA, mB))) return rc;
      if ((rc = run_sim_fft_bin(ctx, nmock, ctx->d_bins))) return rc;
    } else {
      // capture the launch sequence once per (data pointer, nmock, size);
      // the part table is uploaded before capture
      HIPCHK(ctx, hipMemcpyAsync(ctx->d_parts, hp.data(), sizeof(PshDevPart) * nmock,
                                 hipMemcpyHostToDevice, ctx->stream));
      if (!ctx->gexec || ctx->graph_nmock != nmock || ctx->graph_ntotal != ntotal ||
          ctx->graph_data != d_data) {
        drop_graph(ctx);
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        bool prof = ctx->profiling;
        ctx->profiling = false;
        hipError_t e1 = psh_fill_zero_u32((uint32_t *)ctx->d_flags, 4, ctx->stream);
        hipError_t e2 = psh_assign_run(ctx->ag, ctx->d_parts, nmock, ntotal, nmock, ctx->counts,
                                       ctx->starts, ctx->cursors, ctx->records, ctx->rec_cap,
                                       ctx->meshA, mB, ctx->mstride, ctx->scan_tmp,
                                       ctx->d_flags, ctx->stream, nullptr,
                                       ctx->whist_words ? ctx->whist : nullptr);
        int rc2 = (e1 == hipSuccess && e2 == hipSuccess)
                      ? run_sim_fft_bin(ctx, nmock, ctx->d_bins)
                      : PSH_ERR_DEVICE;
        ctx->profiling = prof;
        hipError_t e3 = hipStreamEndCapture(ctx->stream, &ctx->graph);
        if (rc2 || e3 != hipSuccess) {
          drop_graph(ctx);
          return set_err(ctx, PSH_ERR_DEVICE, "graph capture failed");
        }
        hipError_t e4 = fault_graph_now()
                            ? hipErrorOutOfMemory
                            : hipGraphInstantiate(&ctx->gexec, ctx->graph, nullptr, nullptr, 0);
        if (e4 != hipSuccess) {
          ctx->gexec = nullptr;
          drop_graph(ctx);
          return set_err(ctx, PSH_ERR_DEVICE, "graph instantiation failed: %s",
                         hipGetErrorString(e4));
        }
        ctx->graph_nmock = nmock;
        ctx->graph_ntotal = ntotal;
        ctx->graph_data = d_data;
      }
      HIPCHK(ctx, hipGraphLaunch(ctx->gexec, ctx->stream));
    }
    ev_record(ctx, E_END);
    HIPCHK(ctx, hipMemcpyAsync(raw.data(), ctx->d_bins, sizeof(double) * raw.size(),
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(&flags, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (!(flags & 1)) break;
    if (attempt == 3) return set_err(ctx, PSH_ERR_MESH, "assignment record buffer overflow");
    if ((rc = grow_records(ctx))) return rc;
  }
  if (graph) {
    if (ctx->profiling) {
      ctx->ms[3] = ev_ms(ctx, E_START, E_END);
      ctx->ms[0] = ctx->ms[1] = ctx->ms[2] = -1;
    }
  } else {
    ev_collect(ctx);
  }
  if (flags & 4)
    return set_err(ctx, PSH_ERR_CATA, "particle coordinate outside the box [0, BOX_SIZE)");
  if (flags & 2) return set_err(ctx, PSH_ERR_MESH, "inconsistent interlaced cell of a particle");
  const double vol = ctx->L[0] * ctx->L[1] * ctx->L[2];
  const int nmu = ctx->bins.nmu;
  std::vector<double> sums((size_t)ctx->conf.npole * ctx->nbin);
  for (int m = 0; m < nmock; m++) {
    fill_pk_common(ctx, &out[m]);
    const double shot = vol / weight_sum[m];
    const double norm = weight_sum[m] * weight_sum[m] / vol;
    const double *rm = raw.data() + (size_t)m * nacc;
    for (size_t q = 0; q < sums.size(); q++) {   // |mu| cells summed per bin
      double t = 0.0;
      for (int c = 0; c < nmu; c++) t += rm[q * nmu + c];
      sums[q] = t;
    }
    if (out[m].pl[0]) normalise_sim(ctx, sums.data(), shot, norm, out[m].pl[0]);
  }
  return PSH_OK;
}

extern "C" int psh_ctx_define_box(psh_ctx *ctx, const psh_cata *cat) {
  if (!ctx || !cat) return set_err(ctx, PSH_ERR_ARG, "null argument");
  if (cat->num < 1 || cat->num > 2) return set_err(ctx, PSH_ERR_ARG, "invalid catalogue count");
  // bounds on the host (the catalogues are host arrays here)
  double mn[3], mx[3], sd[2], sr[2];
  scan_cat(ctx, cat, mn, mx, sd, sr);
  return def_box(ctx, mn, mx);
}

extern "C" int psh_ctx_run_host(psh_ctx *ctx, psh_cata *cat, psh_pk *out) {
  if (!ctx || !cat || !out) return set_err(ctx, PSH_ERR_ARG, "null argument");
  if (cat->num != ctx->conf.ndata)
    return set_err(ctx, PSH_ERR_ARG, "catalogue count %d != ndata %d", cat->num, ctx->conf.ndata);
  int rc;
  double mn[3], mx[3], swd[2] = {0, 0}, swr[2] = {0, 0};
  scan_cat(ctx, cat, mn, mx, swd, swr);
  if ((rc = def_box(ctx, mn, mx))) return rc;
  if ((rc = ensure_geometry(ctx))) return rc;
  size_t need = 0;
  for (int c = 0; c < cat->num; c++) need += cat->ndata[c] + (ctx->survey ? cat->nrand[c] : 0);
  if (need > ctx->d_cat_cap) {
    if ((rc = dev_alloc(ctx, &ctx->d_cat, need * 4))) return rc;
    ctx->d_cat_cap = need;
  }
  CatDev cd[2] = {};
  size_t off = 0;
  const double vol = ctx->L[0] * ctx->L[1] * ctx->L[2];
  for (int c = 0; c < cat->num; c++) {
    cd[c].data = ctx->d_cat + 4 * off;
    cd[c].nd = cat->ndata[c];
    if ((rc = stage_host(ctx, cat->data[c], cat->ndata[c], off))) return rc;
    off += cat->ndata[c];
    if (ctx->survey) {
      cd[c].rand = ctx->d_cat + 4 * off;
      cd[c].nr = cat->nrand[c];
      if ((rc = stage_host(ctx, cat->rand[c], cat->nrand[c], off))) return rc;
      off += cat->nrand[c];
      cd[c].alpha = cat->alpha[c];
      cd[c].shot = cat->shot[c];
      cd[c].norm = cat->norm[c];
      cd[c].sw_d = swd[c];
      cd[c].sw_r = swr[c];
    } else {
      // shot/norm of simulation boxes (src/genr_mesh.c:904-910)
      const double w = cat->wdata[c];
      cat->shot[c] = vol / w;
      cat->norm[c] = w * w / vol;
      cd[c].shot = cat->shot[c];
      cd[c].norm = cat->norm[c];
    }
    cd[c].wdata = cat->wdata[c];
  }
  return run_catalogues(ctx, cd, cat->num, out);
}

// ===========================================================================
// drop-in layer
// ===========================================================================

extern "C" psh_mesh *psh_genr_mesh(const psh_conf *conf, psh_cata *cat) {
  printf("Generating meshes for FFT ...");
  if (!conf) {
    fprintf(stderr, "\nError: configuration parameters not loaded\n");
    return nullptr;
  }
  if (conf->verbose) printf("\n");
  fflush(stdout);
  if (!cat) {
    fprintf(stderr, "\nError: catalogs not read\n");
    return nullptr;
  }
  psh_mesh *mesh = new (std::nothrow) psh_mesh();
  if (!mesh) return nullptr;
  int rc = psh_ctx_create(conf, 1, &mesh->ctx);
  if (rc) {
    fprintf(stderr, "\nError: failed to initalise the meshes: %s\n", psh_last_error(nullptr));
    delete mesh;
    return nullptr;
  }
  psh_ctx *ctx = mesh->ctx;
  double mn[3], mx[3], swd[2] = {0, 0}, swr[2] = {0, 0};
  if (cat->num < 1 || cat->num > 2) {
    fprintf(stderr, "\nError: invalid number of catalogues: %d\n", cat->num);
    psh_mesh_destroy(mesh);
    return nullptr;
  }
  scan_cat(ctx, cat, mn, mx, swd, swr);
  if ((rc = def_box(ctx, mn, mx))) {
    fprintf(stderr, "\nError: %s\n", ctx->err.c_str());
    psh_mesh_destroy(mesh);
    return nullptr;
  }
  mesh->Ng = ctx->Ng;
  for (int a = 0; a < 3; a++) {
    mesh->bsize[a] = ctx->L[a];
    mesh->bmin[a] = ctx->bmin[a];
  }
  mesh->assign = conf->assign;
  mesh->intlace = conf->intlace;
  mesh->num = cat->num;
  // copy the particles to the device; the assignment runs with powspec so
  // that the device holds one catalogue's meshes at a time
  for (int c = 0; c < cat->num; c++) {
    if (dev_alloc(ctx, &mesh->d_data[c], cat->ndata[c] * 4) ||
        upload_host(ctx, mesh->d_data[c], cat->data[c], sizeof(psh_data) * cat->ndata[c])) {
      fprintf(stderr, "\nError: failed to copy the catalogue to the device\n");
      psh_mesh_destroy(mesh);
      return nullptr;
    }
    mesh->ndata[c] = cat->ndata[c];
    mesh->sw_d[c] = swd[c];
    if (!conf->issim) {
      if (dev_alloc(ctx, &mesh->d_rand[c], cat->nrand[c] * 4) ||
          upload_host(ctx, mesh->d_rand[c], cat->rand[c], sizeof(psh_data) * cat->nrand[c])) {
        fprintf(stderr, "\nError: failed to copy the catalogue to the device\n");
        psh_mesh_destroy(mesh);
        return nullptr;
      }
      mesh->nrand[c] = cat->nrand[c];
      mesh->sw_r[c] = swr[c];
    }
  }
  // shot noise and normalisation of simulation boxes (src/genr_mesh.c:904-910)
  if (conf->issim) {
    for (int c = 0; c < cat->num; c++) {
      double vol = ctx->L[0] * ctx->L[1] * ctx->L[2];
      cat->shot[c] = vol / cat->wdata[c];
      cat->norm[c] = cat->wdata[c] * cat->wdata[c] / vol;
    }
  }
  // release the catalogues like the reference (src/genr_mesh.c:916-922)
  for (int c = 0; c < cat->num; c++) {
    if (cat->data && cat->data[c]) free(cat->data[c]);
    if (cat->rand && cat->rand[c]) free(cat->rand[c]);
  }
  free(cat->data);
  cat->data = nullptr;
  free(cat->rand);
  cat->rand = nullptr;
  printf("\r\x1B[70C[\x1B[32;1mDONE\x1B[0m]\n");
  return mesh;
}

extern "C" psh_pk *psh_powspec(const psh_conf *conf, const psh_cata *cat, psh_mesh *mesh) {
  printf("Evaluating power spectra ...");
  if (!conf) {
    fprintf(stderr, "\nError: configuration parameters not loaded\n");
    return nullptr;
  }
  if (conf->verbose) printf("\n");
  fflush(stdout);
  if (!cat || !mesh || !mesh->ctx) {
    fprintf(stderr, "\nError: %s\n", !cat ? "catalogs not read" : "meshes not generated");
    return nullptr;
  }
  psh_ctx *ctx = mesh->ctx;
  int rc = ensure_geometry(ctx);
  if (rc) {
    fprintf(stderr, "\nError: %s\n", ctx->err.c_str());
    return nullptr;
  }
  const int nb = ctx->nbin, nl = conf->npole;
  psh_pk *pk = (psh_pk *)calloc(1, sizeof(psh_pk));
  if (!pk) return nullptr;
  pk->k = (double *)calloc(nb, sizeof(double));
  pk->kedge = (double *)calloc(nb + 1, sizeof(double));
  pk->km = (double *)calloc(nb, sizeof(double));
  pk->cnt = (uint64_t *)calloc(nb, sizeof(uint64_t));
  pk->lcnt = (double *)calloc((size_t)nb * nl, sizeof(double));
  for (int c = 0; c < 2; c++)
    if (conf->isauto[c]) pk->pl[c] = (double *)calloc((size_t)nb * nl, sizeof(double));
  if (!pk->k || !pk->kedge || !pk->km || !pk->cnt || !pk->lcnt) {
    psh_powspec_destroy(pk);
    return nullptr;
  }
  if (conf->iscross) {
    pk->xpl = (double *)calloc((size_t)nb * nl, sizeof(double));
    if (!pk->xpl) {
      psh_powspec_destroy(pk);
      return nullptr;
    }
  }
  CatDev cd[2] = {};
  for (int c = 0; c < mesh->num; c++) {
    cd[c].data = mesh->d_data[c];
    cd[c].nd = mesh->ndata[c];
    cd[c].rand = mesh->d_rand[c];
    cd[c].nr = mesh->nrand[c];
    cd[c].wdata = cat->wdata[c];
    cd[c].alpha = cat->alpha[c];
    cd[c].shot = cat->shot[c];
    cd[c].norm = cat->norm[c];
    cd[c].sw_d = mesh->sw_d[c];
    cd[c].sw_r = mesh->sw_r[c];
  }
  if ((rc = run_catalogues(ctx, cd, mesh->num, pk))) {
    fprintf(stderr, "\nError: %s\n", ctx->err.c_str());
    psh_powspec_destroy(pk);
    return nullptr;
  }
  fill_pk_common(ctx, pk);
  // release the device meshes and particles; keep the metadata
  // (src/multipole.c:1248-1264)
  for (int c = 0; c < 2; c++) {
    if (mesh->d_data[c]) (void)hipFree(mesh->d_data[c]);
    if (mesh->d_rand[c]) (void)hipFree(mesh->d_rand[c]);
    mesh->d_data[c] = mesh->d_rand[c] = nullptr;
  }
  psh_ctx_destroy(ctx);
  mesh->ctx = nullptr;
  printf("\r\x1B[70C[\x1B[32;1mDONE\x1B[0m]\n");
  return pk;
}

extern "C" void psh_mesh_destroy(psh_mesh *mesh) {
  if (!mesh) return;
  for (int c = 0; c < 2; c++) {
    if (mesh->d_data[c]) (void)hipFree(mesh->d_data[c]);
    if (mesh->d_rand[c]) (void)hipFree(mesh->d_rand[c]);
  }
  if (mesh->ctx) psh_ctx_destroy(mesh->ctx);
  delete mesh;
}

extern "C" void psh_powspec_destroy(psh_pk *pk) {
  if (!pk) return;
  free(pk->k);
  free(pk->kedge);
  free(pk->km);
  free(pk->cnt);
  free(pk->lcnt);
  free(pk->pl[0]);
  free(pk->pl[1]);
  free(pk->xpl);
  free(pk);
}

extern "C" int psh_mesh_info(const psh_mesh *mesh, int *Ng, double bsize[3], double bmin[3],
                             int *assign, int *intlace) {
  if (!mesh) return PSH_ERR_ARG;
  if (Ng) *Ng = mesh->Ng;
  for (int a = 0; a < 3; a++) {
    if (bsize) bsize[a] = mesh->bsize[a];
    if (bmin) bmin[a] = mesh->bmin[a];
  }
  if (assign) *assign = mesh->assign;
  if (intlace) *intlace = mesh->intlace;
  return PSH_OK;
}

// ===========================================================================
// device helpers
// ===========================================================================

extern "C" int psh_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

extern "C" int psh_device_malloc(void **ptr, size_t bytes) {
  if (!ptr) return PSH_ERR_ARG;
  *ptr = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return PSH_ERR_DEVICE;
  return hipMalloc(ptr, bytes) == hipSuccess ? PSH_OK : PSH_ERR_MEMORY;
}

extern "C" int psh_device_free(void *ptr) {
  return hipFree(ptr) == hipSuccess ? PSH_OK : PSH_ERR_DEVICE;
}

extern "C" int psh_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? PSH_OK
                                                                         : PSH_ERR_DEVICE;
}

extern "C" int psh_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? PSH_OK
                                                                         : PSH_ERR_DEVICE;
}

extern "C" int psh_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) == hipSuccess ? PSH_OK
                                                                           : PSH_ERR_DEVICE;
}

extern "C" int psh_device_synchronize(void) {
  return hipDeviceSynchronize() == hipSuccess ? PSH_OK : PSH_ERR_DEVICE;
}

extern "C" int psh_device_arch(char *buf, int len) {
  if (!buf || len < 1) return PSH_ERR_ARG;
  int dev = 0;
  hipDeviceProp_t p;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess)
    return PSH_ERR_DEVICE;
  snprintf(buf, len, "%s", p.gcnArchName);
  return PSH_OK;
}

extern "C" int psh_ctx_run_survey_device(psh_ctx *ctx, const psh_data *d_data, size_t nd,
                                         const psh_data *d_rand, size_t nr, double alpha,
                                         double shot, double norm, psh_pk *out) {
  if (!ctx || !out || (nd && !d_data) || (nr && !d_rand))
    return set_err(ctx, PSH_ERR_ARG, "null argument");
  if (!ctx->survey) return set_err(ctx, PSH_ERR_ARG, "psh_ctx_run_survey_device is for surveys");
  if (ctx->conf.ndata != 1 || ctx->conf.iscross)
    return set_err(ctx, PSH_ERR_ARG, "one catalogue pair per call");
  int rc;
  // box from the coordinate range of data and randoms (src/genr_mesh.c:427-578)
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  double swd = 0, swr = 0;
  if ((rc = bounds_device(ctx, (const double *)d_data, nd, mn, mx, &swd)) ||
      (rc = bounds_device(ctx, (const double *)d_rand, nr, mn, mx, &swr)) ||
      (rc = def_box(ctx, mn, mx)) || (rc = ensure_geometry(ctx)))
    return rc;
  CatDev cd[2] = {};
  cd[0].data = (const double *)d_data;
  cd[0].nd = nd;
  cd[0].rand = (const double *)d_rand;
  cd[0].nr = nr;
  cd[0].alpha = alpha;
  cd[0].shot = shot;
  cd[0].norm = norm;
  cd[0].sw_d = swd;
  cd[0].sw_r = swr;
  return run_catalogues(ctx, cd, 1, out);
}

// ===========================================================================
// deposit tile tuning
// ===========================================================================

extern "C" int psh_ctx_get_tiles(const psh_ctx *ctx, int tile[3]) {
  if (!ctx || !tile) return set_err(nullptr, PSH_ERR_ARG, "null argument");
  if (!ctx->box_set)
    return set_err(const_cast<psh_ctx *>(ctx), PSH_ERR_ARG, "the context has no box yet");
  for (int a = 0; a < 3; a++) tile[a] = ctx->ag.tile[a];
  return PSH_OK;
}

extern "C" int psh_ctx_get_box(const psh_ctx *ctx, double bsize[3], double bmin[3]) {
  if (!ctx || !bsize || !bmin) return set_err(nullptr, PSH_ERR_ARG, "null argument");
  if (!ctx->box_set)
    return set_err(const_cast<psh_ctx *>(ctx), PSH_ERR_ARG, "the context has no box yet");
  for (int a = 0; a < 3; a++) {
    bsize[a] = ctx->L[a];
    bmin[a] = ctx->bmin[a];
  }
  return PSH_OK;
}

extern "C" int psh_fused_tile(int gsize, int intlace, int tile[3]) {
  if (!tile) return 0;
  int t[3];
  if (!psh_zfuse_tile(gsize, intlace, t)) return 0;
  for (int a = 0; a < 3; a++) tile[a] = t[a];
  return 1;
}

extern "C" int psh_ctx_set_tiles(psh_ctx *ctx, const int tile[3]) {
  if (!ctx || !tile) return set_err(ctx, PSH_ERR_ARG, "null argument");
  if (!tiles_ok(tile, ctx->conf.assign, ctx->conf.intlace, ctx->survey ? 8 : 4))
    return set_err(ctx, PSH_ERR_ARG, "invalid tile shape %dx%dx%d", tile[0], tile[1], tile[2]);
  for (int a = 0; a < 3; a++) ctx->tile_req[a] = tile[a];
  if (ctx->box_set) {
    // re-derive the assignment geometry with the new tiles (the box, bins
    // and tables stay); a captured batch graph holds the old tiles
    drop_graph(ctx);
    set_assign_geom(ctx);
  }
  return PSH_OK;
}

namespace {

std::string arch_name() {
  hipDeviceProp_t p;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess)
    return "unknown";
  std::string a = p.gcnArchName;
  const size_t c = a.find(':');
  return c == std::string::npos ? a : a.substr(0, c);
}

// nmock class of a cache entry: 1, or the batch size rounded down to a power of 2
int nmock_class(int nmock) {
  int c = 1;
  while (c * 2 <= nmock) c *= 2;
  return c;
}

}  // namespace

extern "C" int psh_ctx_set_fft_schedule(psh_ctx *ctx, int ysched) {
  if (!ctx) return set_err(nullptr, PSH_ERR_ARG, "null context");
  if (!psh_fft_y_sched_ok(ctx->Ng, ysched))
    return set_err(ctx, PSH_ERR_ARG, "y-pass schedule %d does not apply at GRID_SIZE %d", ysched,
                   ctx->Ng);
  if (ysched != ctx->ysched) drop_graph(ctx);   // a captured graph holds the old launches
  ctx->ysched = ysched;
  return PSH_OK;
}

extern "C" int psh_ctx_get_fft_schedule(const psh_ctx *ctx, int *ysched) {
  if (!ctx || !ysched) return set_err(nullptr, PSH_ERR_ARG, "null argument");
  *ysched = ctx->ysched;
  return PSH_OK;
}

extern "C" int psh_ctx_load_tuning(psh_ctx *ctx, const char *cache_path) {
  if (!ctx || !cache_path) return set_err(ctx, PSH_ERR_ARG, "null argument");
  FILE *fp = fopen(cache_path, "r");
  if (!fp) return set_err(ctx, PSH_ERR_FILE, "cannot open tuning cache %s", cache_path);
  const std::string arch = arch_name();
  const int cls = nmock_class(ctx->nmock_max);
  char line[512];
  int best[3] = {0, 0, 0};
  double best_ms = 1e300;
  int ybest = -1;
  double ybest_ms = 1e300;
  while (fgets(line, sizeof line, fp)) {
    char a[64];
    int Ng, scheme, il, nm, t[3];
    double npart, ms;
    int ys;
    // FFT schedules: "fft arch Ng intlace nmock_class y ysched ms"
    if (sscanf(line, "fft %63s %d %d %d y %d %lf", a, &Ng, &il, &nm, &ys, &ms) == 6) {
      if (arch == a && Ng == ctx->Ng && il == ctx->conf.intlace && nm == cls &&
          psh_fft_y_sched_ok(ctx->Ng, ys) && ms < ybest_ms) {
        ybest_ms = ms;
        ybest = ys;
      }
      continue;
    }
    if (line[0] == '#' ||
        sscanf(line, "%63s %d %d %d %d %lf %d %d %d %lf", a, &Ng, &scheme, &il, &nm, &npart,
               &t[0], &t[1], &t[2], &ms) != 10)
      continue;
    if (arch != a || Ng != ctx->Ng || scheme != ctx->conf.assign ||
        il != ctx->conf.intlace || nm != cls || !tiles_ok(t, scheme, il, ctx->survey ? 8 : 4))
      continue;
    if (ms < best_ms) {   // the latest measurement of equal speed wins ties
      best_ms = ms;
      for (int k = 0; k < 3; k++) best[k] = t[k];
    }
  }
  fclose(fp);
  if (!best[0] && ybest < 0)
    return set_err(ctx, PSH_ERR_ARG, "no tuning entry for this configuration");
  int rc;
  if (best[0] && (rc = psh_ctx_set_tiles(ctx, best))) return rc;
  if (ybest >= 0 && (rc = psh_ctx_set_fft_schedule(ctx, ybest))) return rc;
  return PSH_OK;
}

extern "C" int psh_autotune_fft(const psh_conf *conf, int nmock, const char *cache_path,
                                psh_tune_fft *out) {
  if (!conf || !out || nmock < 1) return set_err(nullptr, PSH_ERR_ARG, "null argument");
  memset(out, 0, sizeof *out);
  // the schedules are timed on the meshes of a box context of the same
  // shape, zeroed (the transform's time does not depend on the values), with
  // the y pass's live-column pruning of that binning
  psh_conf c = *conf;
  c.issim = 1;
  c.ndata = 1;
  c.iscross = 0;
  if (!c.has_bsize || !(c.bsize[0] > 0)) {
    c.has_bsize = 1;
    c.bsize[0] = c.bsize[1] = c.bsize[2] = 1000.0;
  }
  psh_ctx *ctx = nullptr;
  int rc = psh_ctx_create(&c, nmock, &ctx);
  if (rc) return rc;
  const int N = ctx->Ng;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<double> tms(PSH_FFT_Y_SCHEDS, -1.0);
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
      psh_fill_zero_u32((uint32_t *)ctx->meshA, ctx->mstride * nmock, ctx->stream) !=
          hipSuccess)
    rc = set_err(ctx, PSH_ERR_DEVICE, "FFT tuning set-up failed");
  float2 *A = reinterpret_cast<float2 *>(ctx->meshA);
  double best = 1e300;
  for (int ys = 0; ys < PSH_FFT_Y_SCHEDS && !rc; ys++) {
    if (!psh_fft_y_sched_ok(N, ys)) continue;
    double t[5];
    for (int r = -1; r < 5 && !rc; r++) {
      float ms = 0.f;
      if (hipEventRecord(e0, ctx->stream) != hipSuccess ||
          psh_fft_y(N, A, nmock, ctx->mstride, ctx->tw, 0, ctx->stream, ctx->d_ylive, ys) !=
              hipSuccess ||
          hipEventRecord(e1, ctx->stream) != hipSuccess ||
          hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = set_err(ctx, PSH_ERR_DEVICE, "FFT tuning launch failed (schedule %d)", ys);
        break;
      }
      if (r >= 0) t[r] = ms;   // the first launch warms up
    }
    if (rc) break;
    std::sort(t, t + 5);
    tms[ys] = t[2];
    out->ncand++;
    if (ys == 0) out->ms_default = t[2];
    if (t[2] < best) {
      best = t[2];
      out->ysched = ys;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  const std::string arch = arch_name();
  psh_ctx_destroy(ctx);
  if (rc) return set_err(nullptr, rc, "%s", g_last_error.c_str());
  out->ms = best;
  if (cache_path) {
    FILE *fp = fopen(cache_path, "a");
    if (!fp) return set_err(nullptr, PSH_ERR_FILE, "cannot open %s", cache_path);
    fprintf(fp, "fft %s %d %d %d y %d %.5f\n", arch.c_str(), N, conf->intlace,
            nmock_class(nmock), out->ysched, best);
    if (fclose(fp)) return PSH_ERR_FILE;
  }
  return PSH_OK;
}

extern "C" int psh_autotune(const psh_conf *conf, const psh_data *d_data, const size_t *offsets,
                            size_t npart, int nmock, const char *cache_path, psh_tune *out) {
  if (!conf || !out || nmock < 1 || (!d_data && npart == 0) || (d_data && !offsets))
    return PSH_ERR_ARG;
  memset(out, 0, sizeof *out);
  // tiles are geometry only: tune on a uniform periodic box of the same size
  psh_conf c = *conf;
  c.issim = 1;
  c.ndata = 1;
  c.iscross = 0;
  c.isauto[0] = 1;
  c.isauto[1] = 0;
  if (!c.has_bsize) {
    c.has_bsize = 1;
    c.bsize[0] = c.bsize[1] = c.bsize[2] = 1000.0;
  }
  psh_ctx *ctx = nullptr;
  int rc = psh_ctx_create(&c, nmock, &ctx);
  if (rc) return rc;
  std::vector<size_t> off(nmock + 1);
  std::vector<double> ws(nmock);
  psh_data *d = nullptr;
  const psh_data *src = d_data;
  if (d_data) {   // the caller's catalogues (unit weights assumed for timing)
    for (int m = 0; m <= nmock; m++) off[m] = offsets[m];
    for (int m = 0; m < nmock; m++) ws[m] = (double)(off[m + 1] - off[m]);
  } else {
    const size_t ntot = npart * (size_t)nmock;
    if (hipMalloc(&d, sizeof(psh_data) * ntot) != hipSuccess) {
      psh_ctx_destroy(ctx);
      return PSH_ERR_MEMORY;
    }
    rc = psh_gen_uniform_box(d, ntot, c.bsize, 12345);
    for (int m = 0; m <= nmock; m++) off[m] = (size_t)m * npart;
    for (int m = 0; m < nmock; m++) ws[m] = (double)npart;
    src = d;
  }
  const int nb = rc ? 0 : psh_ctx_nbin(ctx);
  std::vector<double> k(nb), ke(nb + 1), km(nb), lc((size_t)nb * c.npole), pl((size_t)nb * c.npole);
  std::vector<uint64_t> cnt(nb);
  std::vector<psh_pk> pk(nmock);
  for (int m = 0; m < nmock; m++) {
    pk[m] = psh_pk{};
    pk[m].k = k.data();
    pk[m].kedge = ke.data();
    pk[m].km = km.data();
    pk[m].cnt = cnt.data();
    pk[m].lcnt = lc.data();
    pk[m].pl[0] = pl.data();
  }
  // candidate shapes: 2^11..2^14 cells, z rows of 16..128 cells, and for
  // boxes of the grid sizes that have it the fused shape (whole z rows, the
  // z transform in the deposit; its time includes that transform)
  const int opts[4] = {4, 8, 16, 32}, zopts[4] = {16, 32, 64, 128};
  int zt[3];
  const bool zf = conf->issim && psh_zfuse_tile(c.gsize, c.intlace, zt);
  double best_ms = 1e300;
  for (int a = 0; a < 4 && !rc; a++)
    for (int b = 0; b < 4 && !rc; b++)
      for (int z = 0; z < 5 && !rc; z++) {
        int t[3] = {opts[a], opts[b], z < 4 ? zopts[z] : 0};
        if (z == 4) {   // the fused shape, once
          if (!zf || a || b) continue;
          for (int q = 0; q < 3; q++) t[q] = zt[q];
        } else {
          const size_t cells = (size_t)t[0] * t[1] * t[2];
          if (cells < 2048 || cells > 16384 || t[0] > t[1] || t[0] > c.gsize ||
              t[1] > c.gsize || t[2] > c.gsize ||
              !tiles_ok(t, c.assign, c.intlace, conf->issim ? 4 : 8))
            continue;
        }
        if ((rc = psh_ctx_set_tiles(ctx, t))) break;
        double times[4];
        for (int r = 0; r < 4 && !rc; r++) {
          const auto t0 = std::chrono::steady_clock::now();
          rc = nmock == 1
                   ? psh_ctx_run_device(ctx, src + off[0], off[1] - off[0], ws[0], &pk[0])
                   : psh_ctx_run_batch(ctx, src, off.data(), ws.data(), nmock, 0, pk.data());
          times[r] = std::chrono::duration<double, std::milli>(
                         std::chrono::steady_clock::now() - t0).count();
        }
        if (rc) break;
        std::sort(times + 1, times + 4);   // the first run warms up
        const double ms = times[2];
        out->ncand++;
        if (ms < best_ms) {
          best_ms = ms;
          for (int q = 0; q < 3; q++) out->tile[q] = t[q];
        }
      }
  if (d) (void)hipFree(d);
  psh_ctx_destroy(ctx);
  if (rc) return rc;
  out->ms = best_ms;
  if (cache_path) {
    FILE *fp = fopen(cache_path, "a");
    if (!fp) return PSH_ERR_FILE;
    fprintf(fp, "%s %d %d %d %d %.0f %d %d %d %.4f\n", arch_name().c_str(), c.gsize, c.assign,
            c.intlace, nmock_class(nmock), (double)(off[nmock] - off[0]) / nmock, out->tile[0],
            out->tile[1], out->tile[2], best_ms);
    if (fclose(fp)) return PSH_ERR_FILE;
  }
  return PSH_OK;
}
