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

__device__ __forceinline__ double legendre(int ell, double x) {
  if (ell == 0) return 1.0;
  double p0 = 1.0, p1 = x;
  for (int n = 1; n < ell; n++) {
    double p2 = ((2 * n + 1) * x * p1 - n * p0) / (n + 1);
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

__device__ __forceinline__ int mode_bin(double kmod2, const PshBins &b,
                                        double &kmod) {
#pragma clang fp contract(off)
  kmod = b.logscale ? 0.5 * log10(kmod2) : sqrt(kmod2);
  if (kmod < b.kedge0 || kmod >= b.kedgeN) return -1;
  int kb = (int)((kmod - b.kedge0) / b.dk);
  if (kb < 0 || kb >= b.nbin) return -1;
  return kb;
}

// One workgroup per bin.  A mode (i, j, k) belongs to bin b when the
// reference's FP64 expression puts it there (mode_bin above); the workgroup
// of bin b visits every (i, j) row, bounds the k_z indices whose |k| can
// fall in [kedge_b, kedge_b+1) -- widened by two indices against rounding --
// and decides each candidate with mode_bin, so every mode is counted by
// exactly one workgroup.  Threads take rows in a fixed stride and sum in
// registers in row order; the block then adds the thread sums in a fixed
// tree.  The results (cnt, km, lcnt) are therefore the same on every run.
constexpr int kGeoThreads = 256;

__global__ void __launch_bounds__(kGeoThreads)
geometry_kernel(int N, PshAxisTables ax, PshBins b, unsigned long long *__restrict__ out_cnt,
                double *__restrict__ out_km, double *__restrict__ out_lcnt) {
#pragma clang fp contract(off)
  __shared__ double red[kGeoThreads];
  __shared__ unsigned long long redc[kGeoThreads];
  const int bin = blockIdx.x;
  const int M = N / 2;
  // |k| range of the bin (log bins: the edges are 0.5 log10 |k|^2)
  double klo = b.kedge0 + b.dk * bin, khi = b.kedge0 + b.dk * (bin + 1);
  if (b.logscale) {
    klo = pow(10.0, klo);
    khi = pow(10.0, khi);
  }
  if (klo < 0) klo = 0;
  const double klo2 = klo * klo, khi2 = khi * khi;
  const double vz = sqrt(ax.ksq[2][1]);   // |k_z| of index 1
  unsigned long long cnt = 0;
  double km = 0.0;
  double lc[PSH_MAX_NL];
  for (int l = 0; l < PSH_MAX_NL; l++) lc[l] = 0.0;
  const size_t rows = (size_t)N * N;
  for (size_t row = threadIdx.x; row < rows; row += kGeoThreads) {
    const int i = (int)(row / N), j = (int)(row - (size_t)i * N);
    const double kmij = ax.ksq[0][i] + ax.ksq[1][j];
    if (kmij > khi2 * (1 + 1e-12) + 1e-300) continue;
    // candidate k_z indices: |k_z| in [sqrt(klo^2 - kxy^2), sqrt(khi^2 - kxy^2))
    const double zlo = klo2 > kmij ? sqrt(klo2 - kmij) / vz : 0.0;
    const double zhi = sqrt(fmax(khi2 - kmij, 0.0)) / vz;
    int k0 = (int)floor(zlo) - 2, k1 = (int)ceil(zhi) + 2;
    if (k0 < 0) k0 = 0;
    if (k1 > M) k1 = M;
    for (int k = k0; k <= k1; k++) {
      double kmod2 = kmij;
      if (k > 0) kmod2 = kmij + ax.ksq[2][k];
      double kmod;
      if (mode_bin(kmod2, b, kmod) != bin) continue;
      const bool single = (k == 0) || (k == M);
      cnt += single ? 1ull : 2ull;
      km += single ? kmod : 2 * kmod;
      if (k == 0 && kmod2 == 0.0) continue;  // DC: no multipole weight
      // mu divides by |k| in both binnings (src/multipole.c:940, 966): with
      // log bins kmod is 0.5 log10 |k|^2, not |k|
      const double kabs = b.logscale ? sqrt(kmod2) : kmod;
      const double mu = (ax.kv[0][i] * b.los[0] + ax.kv[1][j] * b.los[1] +
                         ax.kv[2][k] * b.los[2]) / kabs;
      const double num = single ? 1.0 : 2.0;
      for (int l = 0; l < b.nl; l++) {
        if (!single && (b.poles[l] & 1)) continue;
        lc[l] += num * legendre(b.poles[l], mu);
      }
    }
  }
  // fixed-order tree over the threads
  auto tree = [&](double v) -> double {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = kGeoThreads / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
  };
  redc[threadIdx.x] = cnt;
  __syncthreads();
  for (int h = kGeoThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) redc[threadIdx.x] += redc[threadIdx.x + h];
    __syncthreads();
  }
  const double kms = tree(km);
  if (threadIdx.x == 0) {
    out_cnt[bin] = redc[0];
    out_km[bin] = kms;
  }
  for (int l = 0; l < b.nl; l++) {
    const double t = tree(lc[l]);
    if (threadIdx.x == 0) out_lcnt[l * b.nbin + bin] = t;
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

hipError_t psh_geometry(int Ng, const PshAxisTables &ax, const PshBins &b,
                        unsigned long long *out_cnt, double *out_km, double *out_lcnt,
                        hipStream_t s) {
  if (b.nbin < 1 || b.nl > PSH_MAX_NL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(geometry_kernel, dim3(b.nbin), dim3(kGeoThreads), 0, s, Ng, ax, b,
                     out_cnt, out_km, out_lcnt);
  return hipGetLastError();
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
