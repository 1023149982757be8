// k_gen.hip -- synthetic catalogues on the GPU from counter-based streams
// (philox.h), drawing the same numbers as the host generators in
// powspec_amd/synth.py:
//  * uniform box (positions uniform in [0, L)^3, unit weights);
//  * clustered box: Gaussian field with P(k) = A (k/0.1)^n exp(-(k/kc)^2) on a
//    generator grid (white noise -> FFT -> amplitude -> inverse FFT),
//    lognormal transform, Poisson sampling at the mean density, positions
//    uniform inside the cell, Zel'dovich shift along z (Psi = i k/k^2
//    delta), periodic wrap;
//  * survey sky sample: RA uniform, sin(Dec) uniform, redshift from a
//    truncated Gaussian n(z), FKP weight 1 / (1 + n(z) P0).
#include <float.h>
#include <math.h>

#include <vector>

#include <hip/hip_runtime.h>

#include "philox.h"
#include "powspec_hip.h"
#include "psh_internal.h"

using namespace pshrng;

namespace {

// Largest coordinate kept for a box of size L, so that x * Ng / L < Ng for
// every grid size (the reference's out-of-range write, Q8, never triggers).
__host__ __device__ inline double box_top(double L) { return L * (1.0 - 4e-16) - 1e-9 * L; }

__host__ __device__ inline double wrap_guard(double x, double L) {
  x = fmod(x, L);
  if (x < 0) x += L;
  const double t = box_top(L);
  return x < t ? x : t;
}

__device__ inline double normal(uint64_t seed, uint64_t i, uint32_t stream) {
  const W4 w = draw(seed, i, stream, 0);
  const double u1 = u53(w.v[0], w.v[1]), u2 = u53(w.v[2], w.v[3]);
  return sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * M_PI * u2);   // Box-Muller
}

// Poisson deviate by inversion of the CDF with one uniform; above lam = 600
// (where exp(-lam) nears the FP64 range) the normal approximation with the
// same uniform.
__device__ inline uint32_t poisson_inv(double lam, double u) {
  if (!(lam > 0)) return 0;
  if (lam > 600) {
    const double k = floor(lam + sqrt(lam) * M_SQRT2 * erfinv(2 * u - 1) + 0.5);
    return k > 0 ? (uint32_t)k : 0u;
  }
  double p = exp(-lam), cdf = p;
  uint32_t k = 0;
  const double kcap = lam + 40 * sqrt(lam) + 100;
  while (u >= cdf && k < kcap) {
    k++;
    p *= lam / k;
    cdf += p;
  }
  return k;
}

__global__ void philox_kernel(uint32_t *out, size_t n, uint64_t seed, uint32_t stream,
                              uint32_t d) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const W4 w = draw(seed, i, stream, d);
    for (int j = 0; j < 4; j++) out[4 * i + j] = w.v[j];
  }
}

__global__ void uniform_box_kernel(double *out, size_t n, double L0, double L1, double L2,
                                   uint64_t seed) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const W4 a = draw(seed, i, kUniformPos, 0), b = draw(seed, i, kUniformPos, 1);
    double *p = out + 4 * i;
    p[0] = fmin(u53(a.v[0], a.v[1]) * L0, box_top(L0));
    p[1] = fmin(u53(a.v[2], a.v[3]) * L1, box_top(L1));
    p[2] = fmin(u53(b.v[0], b.v[1]) * L2, box_top(L2));
    p[3] = 1.0;
  }
}

// white noise into a real mesh [ng][ng][pitch]
__global__ void noise_kernel(float *R, int ng, int pitch, uint64_t seed) {
  const size_t ncell = (size_t)ng * ng * ng;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < ncell;
       c += (size_t)gridDim.x * blockDim.x) {
    const size_t row = c / ng;
    R[row * pitch + (c - row * ng)] = (float)normal(seed, c, kNoise);
  }
}

struct PowerLaw {
  double A, n, kc, kf, norm;   // norm = ng^3 / L^3
};

// delta(k) = W(k) sqrt(P(k) ng^3 / L^3); Psi_z(k) = i kz / k^2 delta(k)
__global__ void amplitude_kernel(float2 *K, float2 *Kz, int ng, int H, PowerLaw p) {
  const int M = ng / 2;
  const size_t nm = (size_t)ng * ng * (M + 1);
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < nm;
       q += (size_t)gridDim.x * blockDim.x) {
    const int kz = (int)(q % (M + 1));
    const size_t r = q / (M + 1);
    const int j = (int)(r % ng), i = (int)(r / ng);
    const double kx = p.kf * (i <= M ? i : i - ng), ky = p.kf * (j <= M ? j : j - ng);
    const double kzz = p.kf * kz;
    const double k2 = kx * kx + ky * ky + kzz * kzz;
    double amp = 0, pz = 0;
    if (k2 > 0) {
      const double k = sqrt(k2);
      amp = sqrt(p.A * pow(k / 0.1, p.n) * exp(-(k / p.kc) * (k / p.kc)) * p.norm);
      pz = kzz / k2;
    }
    const size_t pos = ((size_t)i * ng + j) * H + kz;
    const float2 v = K[pos];
    const float2 d = make_float2((float)(v.x * amp), (float)(v.y * amp));
    K[pos] = d;
    Kz[pos] = make_float2((float)(-d.y * pz), (float)(d.x * pz));
  }
}

// FP64 sum and sum of squares of the real field, per block
__global__ void moments_kernel(const float *R, int ng, int pitch, double *part) {
  __shared__ double s1[256], s2[256];
  const size_t ncell = (size_t)ng * ng * ng;
  double a = 0, b = 0;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < ncell;
       c += (size_t)gridDim.x * blockDim.x) {
    const size_t row = c / ng;
    const double v = R[row * pitch + (c - row * ng)];
    a += v;
    b += v * v;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}

__global__ void counts_kernel(const float *R, int ng, int pitch, double mean, double hvar,
                              uint64_t seed, uint32_t *counts) {
  const size_t ncell = (size_t)ng * ng * ng;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < ncell;
       c += (size_t)gridDim.x * blockDim.x) {
    const size_t row = c / ng;
    const double lam = mean * exp((double)R[row * pitch + (c - row * ng)] - hvar);
    const W4 w = draw(seed, c, kPoisson, 0);
    counts[c] = poisson_inv(lam, u53(w.v[0], w.v[1]));
  }
}

__global__ void place_kernel(const uint32_t *starts, const float *Psi, int ng, int pitch,
                             double L, double f, uint64_t seed, double *out) {
  const size_t ncell = (size_t)ng * ng * ng;
  const double cell = L / ng;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < ncell;
       c += (size_t)gridDim.x * blockDim.x) {
    const uint32_t s0 = starts[c], s1 = starts[c + 1];
    if (s0 == s1) continue;
    const size_t row = c / ng;
    const int k = (int)(c - row * ng), j = (int)(row % ng), i = (int)(row / ng);
    const double sz = f * (double)Psi[row * pitch + k];
    for (uint32_t p = s0; p < s1; p++) {
      // keyed by (cell, slot in the cell), so a cell's particles do not
      // depend on the counts of other cells
      const uint64_t key = ((uint64_t)c << 16) | ((p - s0) & 0xffffu);
      const W4 a = draw(seed, key, kPlace, 0), b = draw(seed, key, kPlace, 1);
      double *o = out + 4 * (size_t)p;
      o[0] = wrap_guard((i + u53(a.v[0], a.v[1])) * cell, L);
      o[1] = wrap_guard((j + u53(a.v[2], a.v[3])) * cell, L);
      o[2] = wrap_guard((k + u53(b.v[0], b.v[1])) * cell + sz, L);
      o[3] = 1.0;
    }
  }
}

__device__ inline double phi(double x) { return 0.5 * erfc(-x * M_SQRT1_2); }

__global__ void sky_kernel(double *out, double *nz, size_t n, psh_sky_spec g) {
  const double s0 = sin(g.dec[0] * M_PI / 180), s1 = sin(g.dec[1] * M_PI / 180);
  const double pa = phi((g.zlo - g.zmu) / g.zsigma), pb = phi((g.zhi - g.zmu) / g.zsigma);
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < n;
       o += (size_t)gridDim.x * blockDim.x) {
    const W4 a = draw(g.seed, o, kSky, 0), b = draw(g.seed, o, kSky, 1);
    double *r = out + 4 * o;
    r[0] = g.ra[0] + (g.ra[1] - g.ra[0]) * u53(a.v[0], a.v[1]);
    r[1] = asin(s0 + (s1 - s0) * u53(a.v[2], a.v[3])) * (180 / M_PI);
    const double pz = pa + (pb - pa) * u53(b.v[0], b.v[1]);
    const double z = g.zmu + g.zsigma * M_SQRT2 * erfinv(2 * pz - 1);
    const double dz = (z - g.zmu) / g.zsigma;
    const double nbar = g.nbar0 * exp(-0.5 * dz * dz);
    r[2] = z;
    r[3] = 1.0 / (1.0 + nbar * g.p0);
    if (nz) nz[o] = nbar;
  }
}

unsigned grid_for(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 16384) b = 16384;
  return (unsigned)(b ? b : 1);
}

}  // namespace

extern "C" int psh_gen_philox(uint32_t *d_out, size_t n, uint64_t seed, uint32_t stream,
                              uint32_t d) {
  if (!d_out) return PSH_ERR_ARG;
  hipLaunchKernelGGL(philox_kernel, dim3(grid_for(n)), dim3(256), 0, nullptr, d_out, n, seed,
                     stream, d);
  return hipDeviceSynchronize() == hipSuccess ? PSH_OK : PSH_ERR_DEVICE;
}

extern "C" int psh_gen_uniform_box(psh_data *d_out, size_t n, const double L[3],
                                   uint64_t seed) {
  if (!d_out || !L || !(L[0] > 0 && L[1] > 0 && L[2] > 0)) return PSH_ERR_ARG;
  hipLaunchKernelGGL(uniform_box_kernel, dim3(grid_for(n)), dim3(256), 0, nullptr,
                     (double *)d_out, n, L[0], L[1], L[2], seed);
  return hipDeviceSynchronize() == hipSuccess ? PSH_OK : PSH_ERR_DEVICE;
}

extern "C" int psh_gen_sky(const psh_sky_spec *g, psh_data *d_out, double *d_nz, size_t n) {
  if (!g || !d_out || !(g->zsigma > 0) || !(g->zhi > g->zlo)) return PSH_ERR_ARG;
  hipLaunchKernelGGL(sky_kernel, dim3(grid_for(n)), dim3(256), 0, nullptr, (double *)d_out,
                     d_nz, n, *g);
  return hipDeviceSynchronize() == hipSuccess ? PSH_OK : PSH_ERR_DEVICE;
}

extern "C" int psh_gen_clustered_box(const psh_box_spec *g, psh_data **d_out, size_t *n_out) {
  if (!g || !d_out || !n_out) return PSH_ERR_ARG;
  *d_out = nullptr;
  *n_out = 0;
  const int ng = g->ng;
  if (ng < 4 || ng > 2048 || (ng & (ng - 1)) || !(g->L > 0)) return PSH_ERR_ARG;
  const int M = ng / 2, H = M + PSH_PAD;
  const size_t ncell = (size_t)ng * ng * ng;
  const size_t mfl = (size_t)ng * ng * 2 * H;   // floats per mesh
  hipStream_t s = nullptr;
  float *R = nullptr, *Z = nullptr;
  float2 *tw = nullptr;
  uint32_t *counts = nullptr, *starts = nullptr, *tmp = nullptr;
  double *part = nullptr;
  int rc = PSH_ERR_DEVICE;
  std::vector<double> hp;
  double mean, var, hvar, cellv;
  PshYlmR noy{};
  PowerLaw pl;
  const int pblocks = 1024;
  uint32_t total = 0;
  auto ok = [](hipError_t e) { return e == hipSuccess; };
  if (!ok(hipMalloc(&R, sizeof(float) * mfl)) || !ok(hipMalloc(&Z, sizeof(float) * mfl)) ||
      !ok(hipMalloc(&counts, sizeof(uint32_t) * ncell)) ||
      !ok(hipMalloc(&starts, sizeof(uint32_t) * (ncell + 1))) ||
      !ok(hipMalloc(&tmp, psh_assign_scan_bytes(ncell))) ||
      !ok(hipMalloc(&part, sizeof(double) * 2 * pblocks))) {
    rc = PSH_ERR_MEMORY;
    goto done;
  }
  if (!ok(psh_fft_twiddles(ng, &tw))) goto done;
  // Gaussian field and its z displacement
  hipLaunchKernelGGL(noise_kernel, dim3(grid_for(ncell)), dim3(256), 0, s, R, ng, 2 * H,
                     g->seed);
  if (!ok(psh_fft_z_r2c(ng, R, (float2 *)R, 1, mfl, tw, noy, s)) ||
      !ok(psh_fft_y(ng, (float2 *)R, 1, mfl, tw, 0, s)) ||
      !ok(psh_fft_x(ng, (float2 *)R, (float2 *)R, 1, mfl, tw, 0, s)))
    goto done;
  pl = PowerLaw{g->A, g->nidx, g->kc, 2 * M_PI / g->L,
                (double)ncell / (g->L * g->L * g->L)};
  hipLaunchKernelGGL(amplitude_kernel, dim3(grid_for((size_t)ng * ng * (M + 1))), dim3(256), 0,
                     s, (float2 *)R, (float2 *)Z, ng, H, pl);
  for (float *F : {R, Z}) {
    if (!ok(psh_fft_x(ng, (float2 *)F, (float2 *)F, 1, mfl, tw, 1, s)) ||
        !ok(psh_fft_y(ng, (float2 *)F, 1, mfl, tw, 1, s)) ||
        !ok(psh_fft_z_c2r(ng, (float2 *)F, 1, mfl, tw, (float)(1.0 / (double)ncell), s)))
      goto done;
  }
  // variance of delta_G for the lognormal transform
  hipLaunchKernelGGL(moments_kernel, dim3(pblocks), dim3(256), 0, s, R, ng, 2 * H, part);
  hp.resize(2 * pblocks);
  if (!ok(hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost,
                         s)) ||
      !ok(hipStreamSynchronize(s)))
    goto done;
  {
    double s1 = 0, s2 = 0;
    for (int b = 0; b < pblocks; b++) {
      s1 += hp[2 * b];
      s2 += hp[2 * b + 1];
    }
    const double mu = s1 / (double)ncell;
    var = s2 / (double)ncell - mu * mu;
  }
  hvar = 0.5 * var;
  cellv = g->L / ng;
  mean = g->npart > 0 ? g->npart / (double)ncell : g->nbar * cellv * cellv * cellv;
  hipLaunchKernelGGL(counts_kernel, dim3(grid_for(ncell)), dim3(256), 0, s, R, ng, 2 * H, mean,
                     hvar, g->seed, counts);
  if (!ok(psh_exclusive_scan_u32(counts, starts, tmp, ncell, s)) ||
      !ok(hipMemcpyAsync(&total, starts + ncell, sizeof(uint32_t), hipMemcpyDeviceToHost, s)) ||
      !ok(hipStreamSynchronize(s)))
    goto done;
  if (!ok(hipMalloc((void **)d_out, sizeof(psh_data) * (total ? total : 1)))) {
    rc = PSH_ERR_MEMORY;
    goto done;
  }
  hipLaunchKernelGGL(place_kernel, dim3(grid_for(ncell)), dim3(256), 0, s, starts, Z, ng, 2 * H,
                     g->L, g->f, g->seed, (double *)*d_out);
  if (!ok(hipStreamSynchronize(s))) goto done;
  *n_out = total;
  rc = PSH_OK;
done:
  if (rc != PSH_OK && *d_out) {
    (void)hipFree(*d_out);
    *d_out = nullptr;
  }
  for (void *p : {(void *)R, (void *)Z, (void *)tw, (void *)counts, (void *)starts, (void *)tmp,
                  (void *)part})
    if (p) (void)hipFree(p);
  return rc;
}
