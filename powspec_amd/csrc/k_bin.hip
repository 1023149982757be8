// k_bin.hip -- k-bin geometry, deterministic reductions and coordinate
// bounds on gfx950.
//
// psh_geometry replaces the mode counting of powspec_precomp
// (src/multipole.c:111-257) and the lcnt accumulation of count_mode_sim_lin
// (src/multipole.c:782-892), which depend on the box and bins only: it is run
// once per geometry and cached.  The bin index of each mode is computed in
// FP64 in the reference's operation order with contraction disabled, so the
// integer counts are bit-identical to the reference's.
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

__global__ void __launch_bounds__(256)
geometry_kernel(int N, PshAxisTables ax, PshBins b,
                unsigned long long *__restrict__ part_cnt,
                double *__restrict__ part_km, double *__restrict__ part_lcnt) {
#pragma clang fp contract(off)
  extern __shared__ double sh[];
  const int nb = b.nbin;
  unsigned long long *cnt = reinterpret_cast<unsigned long long *>(sh);
  double *km = sh + nb;
  double *lc = sh + 2 * nb;
  for (int i = threadIdx.x; i < (2 + b.nl) * nb; i += blockDim.x) sh[i] = 0.0;
  __syncthreads();
  const int M = N / 2;
  const size_t rows = (size_t)N * N;
  for (size_t row = blockIdx.x; row < rows; row += gridDim.x) {
    const int i = (int)(row / N), j = (int)(row - (size_t)i * N);
    const double kmij = ax.ksq[0][i] + ax.ksq[1][j];
    for (int k = threadIdx.x; k <= M; k += blockDim.x) {
      double kmod2 = kmij;
      if (k > 0) kmod2 = kmij + ax.ksq[2][k];
      double kmod;
      int kb = mode_bin(kmod2, b, kmod);
      if (kb < 0) continue;
      const bool single = (k == 0) || (k == M);
      atomicAdd(&cnt[kb], single ? 1ull : 2ull);
      atomicAdd(&km[kb], single ? kmod : 2 * kmod);
      if (k == 0 && kmod2 == 0.0) continue;  // DC: no multipole weight
      // mu divides by |k| in both binnings (src/multipole.c:940, 966): with
      // log bins kmod is 0.5 log10 |k|^2, not |k|
      const double kabs = b.logscale ? sqrt(kmod2) : kmod;
      const double mu = (ax.kv[0][i] * b.los[0] + ax.kv[1][j] * b.los[1] +
                         ax.kv[2][k] * b.los[2]) / kabs;
      const double num = single ? 1.0 : 2.0;
      for (int l = 0; l < b.nl; l++) {
        if (!single && (b.poles[l] & 1)) continue;
        atomicAdd(&lc[l * nb + kb], num * legendre(b.poles[l], mu));
      }
    }
  }
  __syncthreads();
  const size_t o = (size_t)blockIdx.x;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    part_cnt[o * nb + i] = cnt[i];
    part_km[o * nb + i] = km[i];
  }
  for (int i = threadIdx.x; i < b.nl * nb; i += blockDim.x)
    part_lcnt[o * b.nl * nb + i] = lc[i];
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
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
       p += (size_t)gridDim.x * blockDim.x) {
    const double2 a = reinterpret_cast<const double2 *>(xyzw + 4 * p)[0];
    const double z = xyzw[4 * p + 2];
    mn[0] = fmin(mn[0], a.x);
    mx[0] = fmax(mx[0], a.x);
    mn[1] = fmin(mn[1], a.y);
    mx[1] = fmax(mx[1], a.y);
    mn[2] = fmin(mn[2], z);
    mx[2] = fmax(mx[2], z);
  }
  __shared__ double sh[6][256];
  for (int a = 0; a < 3; a++) {
    sh[a][threadIdx.x] = mn[a];
    sh[3 + a][threadIdx.x] = mx[a];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      for (int a = 0; a < 3; a++) {
        sh[a][threadIdx.x] = fmin(sh[a][threadIdx.x], sh[a][threadIdx.x + s]);
        sh[3 + a][threadIdx.x] =
            fmax(sh[3 + a][threadIdx.x], sh[3 + a][threadIdx.x + s]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

}  // namespace

hipError_t psh_geometry(int Ng, const PshAxisTables &ax, const PshBins &b,
                        unsigned long long *part_cnt, double *part_km,
                        double *part_lcnt, int grid, unsigned long long *out_cnt,
                        double *out_km, double *out_lcnt, hipStream_t s) {
  const size_t shmem = sizeof(double) * (size_t)(2 + b.nl) * b.nbin;
  if (shmem > 160 * 1024) return hipErrorInvalidValue;
  if (shmem > 64 * 1024)
    (void)hipFuncSetAttribute((const void *)geometry_kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
  hipLaunchKernelGGL(geometry_kernel, dim3(grid), dim3(256), shmem, s, Ng, ax, b,
                     part_cnt, part_km, part_lcnt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if ((e = launch_reduce<unsigned long long>(part_cnt, 1, grid, b.nbin, out_cnt, s)))
    return e;
  if ((e = launch_reduce<double>(part_km, 1, grid, b.nbin, out_km, s))) return e;
  return launch_reduce<double>(part_lcnt, 1, grid, b.nl * b.nbin, out_lcnt, s);
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
