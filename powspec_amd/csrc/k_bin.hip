// k_bin.hip -- k-bin geometry, deterministic reductions and coordinate
// bounds on gfx950.
//
// psh_geometry replaces the mode counting of powspec_precomp
// (src/multipole.c:111-257) and the lcnt accumulation of count_mode_sim_lin
// (src/multipole.c:782-892), which depend on the box and bins only: it is run
// once per geometry and cached.  The bin index of each mode is computed in
// FP64 in the reference's operation order with contraction disabled, so the
// integer counts are bit-identical to the reference's; the FP64 sums are
// added in a fixed order, so they are the same on every run.
#include "psh_internal.h"

#include <float.h>

namespace {


__device__ __forceinline__ int mode_bin(double kmod2, const PshBins &b,
                                        double &kmod) {
#pragma clang fp contract(off)
  kmod = b.logscale ? 0.5 * log10(kmod2) : sqrt(kmod2);
  if (kmod < b.kedge0 || kmod >= b.kedgeN) return -1;
  int kb = (int)((kmod - b.kedge0) / b.dk);
  if (kb < 0 || kb >= b.nbin) return -1;
  return kb;
}

// Mode-parallel geometry with order-independent sums.  A work item is a run
// of kGeoRun consecutive k_z of one (i, j) row; along a run |k| only grows,
// so its modes fall in consecutive bins and a thread sums each bin's modes
// in registers (FP64, in k_z order) and adds the run total to the
// workgroup's LDS accumulators only when the bin changes.  Those additions
// are integer: cnt exactly, km and lcnt as two-word fixed point (x 2^s split
// into floor(x 2^s) and the remainder's 32 fractional bits, both summed as
// 64-bit integers, so every sum is exact and independent of the order in
// which threads arrive).  Each workgroup converts its sums to FP64 and
// writes them as its row of the partials, which the fixed-order reduction
// adds up: every output is the same on every run.  Bins beyond the LDS go
// in windows of nbw bins, one launch each.
constexpr int kGeoThreads = 256;
constexpr int kGeoRun = 32;
constexpr int kGeoGrid = 1024;
constexpr int kGeoMaxWindow = 1024;

// L_l(mu) for l <= 6 in the reference's FP64 expressions
// (math/legpoly.h:47-61; 3/14 is its 0x1.b6db6db6db6dbp-3)
__device__ __forceinline__ double legendre_h(int ell, double x) {
#pragma clang fp contract(off)
  const double x2 = x * x;
  switch (ell) {
    case 0: return 1.0;
    case 1: return x;
    case 2: return 1.5 * x2 - 0.5;
    case 3: return 2.5 * x * (x2 - 0.6);
    case 4: return 4.375 * x2 * x2 - 3.75 * x2 + 0.375;
    case 5: return 8.75 * x * (0.9 * x2 * x2 - x2 + 3.0 / 14.0);
    default: return 14.4375 * x2 * x2 * x2 - 19.6875 * x2 * x2 + 6.5625 * x2 - 0.3125;
  }
}

struct GeoArgs {
  int N;
  PshAxisTables ax;
  PshBins b;
  int kb0, nbw;
  double skm, slc;   // fixed-point scales 2^s of km and lcnt
  double kskip2;     // rows with k_x^2 + k_y^2 above this hold no binned mode
};

__device__ __forceinline__ void fixed_add(long long *hi, long long *lo, double x, double sc) {
  const double y = x * sc;                        // exact (power of two)
  const double h = floor(y);                      // exact
  const double r = (y - h) * 4294967296.0;       // exact, in [0, 2^32)
  atomicAdd(reinterpret_cast<unsigned long long *>(hi), (unsigned long long)(long long)h);
  atomicAdd(reinterpret_cast<unsigned long long *>(lo), (unsigned long long)(long long)rint(r));
}

__device__ __forceinline__ double fixed_value(long long hi, long long lo, double sc) {
  return (double)hi / sc + (double)lo / (sc * 4294967296.0);
}

__global__ void __launch_bounds__(kGeoThreads)
geometry_kernel(GeoArgs g, unsigned long long *__restrict__ part_cnt,
                double *__restrict__ part_km, double *__restrict__ part_lc) {
#pragma clang fp contract(off)
  extern __shared__ long long gsh[];
  const PshBins &b = g.b;
  const int nbw = g.nbw, nl = b.nl;
  unsigned long long *cnt = reinterpret_cast<unsigned long long *>(gsh);
  long long *kmh = gsh + nbw, *kml = gsh + 2 * nbw;
  long long *lch = gsh + 3 * nbw, *lcl = gsh + (3 + nl) * nbw;
  for (int i = threadIdx.x; i < (3 + 2 * nl) * nbw; i += kGeoThreads) gsh[i] = 0;
  __syncthreads();
  const int N = g.N, M = N / 2;
  const unsigned nseg = (M + kGeoRun) / kGeoRun;   // runs per row: ceil((M + 1) / kGeoRun)
  // rows (i, j) with 0 <= i, j <= M; each stands for its mirrors N - i and
  // N - j as well (0 < i, j < M), which share |k|^2 exactly (ksq of the
  // signed index), so the bin, |k| and the run structure are computed once
  const unsigned nq = (unsigned)(M + 1);
  const unsigned items = nq * nq * nseg;
  const unsigned stride = gridDim.x * kGeoThreads;
  int cur = -1;
  unsigned long long rc = 0;
  double rkm = 0.0, rlc[PSH_MAX_NL];
  auto flush = [&]() {
    const int w = cur - g.kb0;
    if (cur >= 0 && w >= 0 && w < nbw) {
      atomicAdd(&cnt[w], rc);
      fixed_add(&kmh[w], &kml[w], rkm, g.skm);
      for (int l = 0; l < nl; l++) fixed_add(&lch[l * nbw + w], &lcl[l * nbw + w], rlc[l], g.slc);
    }
    rc = 0;
    rkm = 0.0;
    for (int l = 0; l < PSH_MAX_NL; l++) rlc[l] = 0.0;
  };
  for (int l = 0; l < PSH_MAX_NL; l++) rlc[l] = 0.0;
  for (unsigned it = blockIdx.x * kGeoThreads + threadIdx.x; it < items; it += stride) {
    const unsigned row = it / nseg;
    const int seg = (int)(it - row * nseg);
    const int i = (int)(row / nq), j = (int)(row - (unsigned)i * nq);
    const double kmij = g.ax.ksq[0][i] + g.ax.ksq[1][j];
    if (kmij > g.kskip2) continue;
    // mirrors: x -> -x when 0 < i < M, y -> -y when 0 < j < M
    const int mx = (i > 0 && i < M) ? 2 : 1, my = (j > 0 && j < M) ? 2 : 1;
    const unsigned nmir = (unsigned)(mx * my);
    // the reference's mu numerator, ((k_x l_x + k_y l_y) + k_z l_z), of a
    // mirrored mode is the same expression with the mirrored products, and
    // the negation of a product is exact
    const double px = g.ax.kv[0][i] * b.los[0], py = g.ax.kv[1][j] * b.los[1];
    // mirrors along an axis the line of sight does not follow share mu:
    // one Legendre evaluation stands for them (weight 2, exact)
    const int ux = b.los[0] == 0.0 ? 1 : mx, uy = b.los[1] == 0.0 ? 1 : my;
    const double wmir = (double)((mx / ux) * (my / uy));
    const int k0 = seg * kGeoRun, k1 = min(k0 + kGeoRun - 1, M);
    cur = -1;
    for (int k = k0; k <= k1; k++) {
      double kmod2 = kmij;
      if (k > 0) kmod2 = kmij + g.ax.ksq[2][k];
      double kmod;
      const int kb = mode_bin(kmod2, b, kmod);
      if (kb != cur) {
        flush();
        cur = kb;
      }
      if (kb < 0) continue;
      const bool single = (k == 0) || (k == M);
      rc += (single ? 1ull : 2ull) * nmir;
      rkm += (single ? kmod : 2 * kmod) * (double)nmir;   // nmir = 1, 2, 4: exact
      if (k == 0 && kmod2 == 0.0) continue;  // DC: no multipole weight
      // mu divides by |k| in both binnings (src/multipole.c:940, 966): with
      // log bins kmod is 0.5 log10 |k|^2, not |k|
      const double kabs = b.logscale ? sqrt(kmod2) : kmod;
      const double pz = g.ax.kv[2][k] * b.los[2];
      const double num = (single ? 1.0 : 2.0) * wmir;
      for (int sx = 0; sx < ux; sx++)
        for (int sy = 0; sy < uy; sy++) {
          const double mu = ((sx ? -px : px) + (sy ? -py : py) + pz) / kabs;
          for (int l = 0; l < nl; l++) {
            if (!single && (b.poles[l] & 1)) continue;
            rlc[l] += num * legendre_h(b.poles[l], mu);
          }
        }
    }
    flush();
  }
  __syncthreads();
  const size_t o = (size_t)blockIdx.x * nbw;
  for (int w = threadIdx.x; w < nbw; w += kGeoThreads) {
    part_cnt[o + w] = cnt[w];
    part_km[o + w] = fixed_value(kmh[w], kml[w], g.skm);
    for (int l = 0; l < nl; l++)
      part_lc[(size_t)l * gridDim.x * nbw + o + w] =
          fixed_value(lch[l * nbw + w], lcl[l * nbw + w], g.slc);
  }
}

// out[m][a] = sum over g of part[m][g][a].  A block takes 16 accumulators;
// its 16 slices of threads each sum a contiguous range of g in order, and the
// slice sums are added in slice order: the result does not depend on timing.
constexpr int kRedAcc = 16, kRedSlices = 16;

template <class T>
__global__ void __launch_bounds__(kRedAcc * kRedSlices)
reduce_kernel(const T *__restrict__ part, int grid, int nacc, T *__restrict__ out) {
  __shared__ T sh[kRedSlices][kRedAcc];
  const int l = threadIdx.x % kRedAcc, slice = threadIdx.x / kRedAcc;
  const int a = blockIdx.x * kRedAcc + l;
  const int m = blockIdx.y;
  T s = 0;
  if (a < nacc) {
    const T *p = part + (size_t)m * grid * nacc + a;
    const int g0 = (int)((long)slice * grid / kRedSlices);
    const int g1 = (int)((long)(slice + 1) * grid / kRedSlices);
#pragma unroll 8
    for (int g = g0; g < g1; g++) s += p[(size_t)g * nacc];
  }
  sh[slice][l] = s;
  __syncthreads();
  if (slice == 0 && a < nacc) {
    T t = sh[0][l];
    for (int q = 1; q < kRedSlices; q++) t += sh[q][l];
    out[(size_t)m * nacc + a] = t;
  }
}

template <class T>
hipError_t launch_reduce(const T *part, int nmock, int grid, int nacc, T *out,
                         hipStream_t s) {
  const dim3 blocks((unsigned)((nacc + kRedAcc - 1) / kRedAcc), (unsigned)nmock);
  hipLaunchKernelGGL(reduce_kernel<T>, blocks, dim3(kRedAcc * kRedSlices), 0, s, part,
                     grid, nacc, out);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256)
bounds_kernel(const double *__restrict__ xyzw, size_t n,
              double *__restrict__ partial) {
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
  double mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  double ws = 0.0;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
       p += (size_t)gridDim.x * blockDim.x) {
    const double2 a = reinterpret_cast<const double2 *>(xyzw + 4 * p)[0];
    const double2 zw = reinterpret_cast<const double2 *>(xyzw + 4 * p)[1];
    const double z = zw.x;
    ws += zw.y;
    mn[0] = fmin(mn[0], a.x);
    mx[0] = fmax(mx[0], a.x);
    mn[1] = fmin(mn[1], a.y);
    mx[1] = fmax(mx[1], a.y);
    mn[2] = fmin(mn[2], z);
    mx[2] = fmax(mx[2], z);
  }
  __shared__ double sh[7][256];
  for (int a = 0; a < 3; a++) {
    sh[a][threadIdx.x] = mn[a];
    sh[3 + a][threadIdx.x] = mx[a];
  }
  sh[6][threadIdx.x] = ws;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      for (int a = 0; a < 3; a++) {
        sh[a][threadIdx.x] = fmin(sh[a][threadIdx.x], sh[a][threadIdx.x + s]);
        sh[3 + a][threadIdx.x] =
            fmax(sh[3 + a][threadIdx.x], sh[3 + a][threadIdx.x + s]);
      }
      sh[6][threadIdx.x] += sh[6][threadIdx.x + s];   // fixed tree: same sum every run
    }
    __syncthreads();
  }
  static_assert(PSH_BOUNDS_STRIDE == 7, "bounds partial layout");
  if (threadIdx.x < 7) partial[blockIdx.x * PSH_BOUNDS_STRIDE + threadIdx.x] = sh[threadIdx.x][0];
}

}  // namespace

size_t psh_geometry_scratch_bytes(const PshBins &b) {
  const int nbw = b.nbin < kGeoMaxWindow ? b.nbin : kGeoMaxWindow;
  return sizeof(double) * (size_t)kGeoGrid * nbw * (2 + b.nl);
}

hipError_t psh_geometry(int Ng, const PshAxisTables &ax, const PshBins &b,
                        unsigned long long *out_cnt, double *out_km, double *out_lcnt,
                        void *scratch, hipStream_t s) {
  if (b.nbin < 1 || b.nl > PSH_MAX_NL) return hipErrorInvalidValue;
  // window: LDS of (3 + 2 nl) 64-bit words per bin, at most kGeoMaxWindow bins
  const size_t per_bin = sizeof(long long) * (3 + 2 * (size_t)b.nl);
  int nbw = (int)((size_t)64 * 1024 / per_bin);
  if (nbw > kGeoMaxWindow) nbw = kGeoMaxWindow;
  if (nbw > b.nbin) nbw = b.nbin;
  GeoArgs g{};
  g.N = Ng;
  g.ax = ax;
  g.b = b;
  // fixed-point scales: a workgroup takes at most `modes` modes, each adding
  // at most 2 max|kmod| to km and 2 to each lcnt (|L_l| <= 1); 2^s times the
  // largest possible sum stays below 2^61
  const int M = Ng / 2;
  const double items = (double)(M + 1) * (M + 1) * ((M + kGeoRun) / kGeoRun);
  const double per_thread = ceil(items / ((double)kGeoGrid * kGeoThreads));
  const double modes = per_thread * kGeoThreads * kGeoRun * 4;   // up to 4 mirrors per row
  const double kbig = fmax(fmax(fabs(b.kedge0), fabs(b.kedgeN)), 1e-300);
  int e;
  (void)frexp(2.0 * kbig * modes, &e);
  g.skm = ldexp(1.0, 61 - e);
  (void)frexp(2.0 * modes, &e);
  g.slc = ldexp(1.0, 61 - e);
  // rows whose k_x^2 + k_y^2 alone reaches the last edge have no binned mode
  const double top = b.logscale ? pow(10.0, 2.0 * b.kedgeN) : b.kedgeN * b.kedgeN;
  g.kskip2 = top * (1.0 + 1e-9) + 1e-300;
  unsigned long long *pc = static_cast<unsigned long long *>(scratch);
  double *pk = reinterpret_cast<double *>(pc + (size_t)kGeoGrid * nbw);
  double *pl = pk + (size_t)kGeoGrid * nbw;
  hipError_t err;
  for (int kb0 = 0; kb0 < b.nbin; kb0 += nbw) {
    g.kb0 = kb0;
    g.nbw = b.nbin - kb0 < nbw ? b.nbin - kb0 : nbw;
    const size_t lds = per_bin * g.nbw;
    hipLaunchKernelGGL(geometry_kernel, dim3(kGeoGrid), dim3(kGeoThreads), lds, s, g, pc, pk, pl);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    if ((err = launch_reduce<unsigned long long>(pc, 1, kGeoGrid, g.nbw, out_cnt + kb0, s)))
      return err;
    if ((err = launch_reduce<double>(pk, 1, kGeoGrid, g.nbw, out_km + kb0, s))) return err;
    for (int l = 0; l < b.nl; l++)
      if ((err = launch_reduce<double>(pl + (size_t)l * kGeoGrid * g.nbw, 1, kGeoGrid, g.nbw,
                                       out_lcnt + (size_t)l * b.nbin + kb0, s)))
        return err;
  }
  return hipSuccess;
}

hipError_t psh_reduce_partials(const double *partials, int nmock, int grid, int nacc,
                               double *out, hipStream_t s) {
  return launch_reduce<double>(partials, nmock, grid, nacc, out, s);
}

hipError_t psh_coord_bounds(const double *xyzw, size_t n, double *partial,
                            int blocks, hipStream_t s) {
  hipLaunchKernelGGL(bounds_kernel, dim3(blocks), dim3(256), 0, s, xyzw, n, partial);
  return hipGetLastError();
}
