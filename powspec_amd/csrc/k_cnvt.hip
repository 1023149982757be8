// k_cnvt.hip -- survey coordinate conversion on gfx950 (src/cnvt_coord.c):
// (RA, Dec, redshift) -> comoving Cartesian coordinates, FP64.
#include <float.h>
#include <math.h>
#include <string.h>

#include <vector>

#include <hip/hip_runtime.h>

#include "powspec_hip.h"
#include "psh_internal.h"

namespace {

constexpr int kMinOrder = 4, kMaxOrder = 32;   // math/legauss.h:40-41
constexpr int kNumZ = 128;                     // POWSPEC_INT_NUM_ZSP (src/define.h:93)
constexpr double kCH = 299792.458 * 0.01;      // c / (100 km/s/Mpc), Mpc/h
constexpr double kDeg = M_PI / 180;

// Gauss-Legendre rule of order n: the n/2 positive abscissas with weights and
// the weight of x = 0 for odd n, from Newton's iteration on P_n in long
// double (P_n, P_n' by the three-term recurrence).
struct Rule {
  double x[kMaxOrder / 2 + 1], w[kMaxOrder / 2 + 1], w0;
};

Rule legauss_rule(int n) {
  Rule r{};
  auto legendre = [n](long double t, long double &pn, long double &pn1) {
    long double a = 1, b = t;
    for (int j = 2; j <= n; j++) {
      long double c = ((2 * j - 1) * t * b - (j - 1) * a) / j;
      a = b;
      b = c;
    }
    pn = b;
    pn1 = a;
  };
  for (int k = 0; k < n / 2; k++) {
    long double t = cosl(M_PI * (k + 0.75L) / (n + 0.5L)), pn, pn1;
    for (int it = 0; it < 100; it++) {
      legendre(t, pn, pn1);
      long double dt = pn / (n * (t * pn - pn1) / (t * t - 1));
      t -= dt;
      if (fabsl(dt) < 1e-19L) break;
    }
    legendre(t, pn, pn1);
    long double dp = n * (t * pn - pn1) / (t * t - 1);
    r.x[k] = (double)t;
    r.w[k] = (double)(2 / ((1 - t * t) * dp * dp));
  }
  if (n & 1) {
    long double pn, pn1;
    legendre(0, pn, pn1);
    long double dp = n * pn1;
    r.w0 = (double)(2 / (dp * dp));
  }
  return r;
}

struct Cosmo {
  double om, ol, ok, widx;
};

// c/100/E(z) with the terms in the reference's order (src/cnvt_coord.c:241-252)
__host__ __device__ inline double integrand(const Cosmo &c, double z) {
  double z1 = z + 1;
  double z2 = z1 * z1;
  double d = c.om * z2 * z1;
  if (c.ok) d += c.ok * z2;
  if (c.widx) d += c.ol * pow(z1, c.widx);
  else d += c.ol;
  return kCH / sqrt(d);
}

struct RuleDev {
  int n;
  double x[kMaxOrder / 2 + 1], w[kMaxOrder / 2 + 1], w0;
};

__host__ __device__ inline double cmvdist(const Cosmo &c, const RuleDev &r, double z) {
  double zp = z * 0.5, sum = 0;
  for (int k = 0; k < r.n / 2; k++)
    sum += r.w[k] * (integrand(c, zp * (1 + r.x[k])) + integrand(c, zp * (1 - r.x[k])));
  if (r.n & 1) sum += r.w0 * integrand(c, zp);
  return sum * zp;
}

RuleDev rule_dev(int n) {
  Rule r = legauss_rule(n);
  RuleDev d{};
  d.n = n;
  memcpy(d.x, r.x, sizeof d.x);
  memcpy(d.w, r.w, sizeof d.w);
  d.w0 = r.w0;
  return d;
}

__device__ inline void to_cartesian(double *p, double dist) {
  const double ra = p[0] * kDeg, dec = p[1] * kDeg;
  p[0] = dist * cos(dec) * cos(ra);
  p[1] = dist * cos(dec) * sin(ra);
  p[2] = dist * sin(dec);
}

__global__ void __launch_bounds__(256) integr_kernel(double *d, size_t n, Cosmo c, RuleDev r) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double *p = d + 4 * i;
    to_cartesian(p, cmvdist(c, r, p[2]));
  }
}

// Natural cubic spline with the reference's interval search
// (src/cnvt_coord.c:97-111): HUGE_VAL outside [zs[0], zs[n-1]).
__global__ void __launch_bounds__(256) interp_kernel(double *d, size_t n, const double *zs,
                                                     const double *ds, const double *ypp,
                                                     size_t nsp) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double *p = d + 4 * i;
    const double z = p[2];
    double dist = HUGE_VAL;
    if (z >= zs[0] && z < zs[nsp - 1]) {
      size_t lo = 0, hi = nsp - 1;
      while (hi - lo > 1) {
        const size_t mid = (lo + hi) >> 1;
        if (zs[mid] <= z) lo = mid;
        else hi = mid;
      }
      const double b = zs[lo + 1] - zs[lo], a = z - zs[lo], e = zs[lo + 1] - z;
      dist = (a * ds[lo + 1] + e * ds[lo] +
              ((a * a - b * b) * a * ypp[lo + 1] + (e * e - b * b) * e * ypp[lo]) / 6) / b;
    }
    to_cartesian(p, dist);
  }
}

// Second derivatives of the natural spline (zero at both ends).
std::vector<double> spline_ypp(const double *x, const double *y, size_t n) {
  std::vector<double> ypp(n, 0.0), c(n, 0.0);
  for (size_t j = 1; j + 1 < n; j++) {
    const double h0 = x[j] - x[j - 1], h1 = x[j + 1] - x[j];
    const double rhs = 6 * ((y[j + 1] - y[j]) / h1 - (y[j] - y[j - 1]) / h0);
    const double piv = 2 * (h0 + h1) - h0 * c[j - 1];
    ypp[j] = (rhs - h0 * ypp[j - 1]) / piv;
    c[j] = h1 / piv;
  }
  for (size_t j = n - 1; j-- > 0;) ypp[j] -= c[j] * ypp[j + 1];
  return ypp;
}

// Order by the reference's convergence test on kNumZ redshifts spanning
// [zmin, zmax] (src/cnvt_coord.c:306-341): orders are tried from kMinOrder
// up; the test reaches order kMaxOrder + 1 before it gives up (the reference
// reads that rule past the end of its table; here it is computed), -1 when
// that does not converge either.
int choose_order(const Cosmo &c, double err, double zmin, double zmax) {
  std::vector<RuleDev> rules;
  for (int n = kMinOrder; n <= kMaxOrder + 1; n++) rules.push_back(rule_dev(n));
  int order = 0;
  for (int i = 0; i < kNumZ; i++) {
    const double z = zmin + i * (zmax - zmin) / (kNumZ - 1);
    double oint, nint = 0;
    int n = kMinOrder - 1;
    do {
      if (n > kMaxOrder) return -1;
      oint = nint;
      ++n;
      nint = cmvdist(c, rules[n - kMinOrder], z);
    } while (fabs(nint - oint) > nint * err);
    if (order < n) order = n;
  }
  return order;
}

unsigned blocks_for(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  return (unsigned)(b ? b : 1);
}

}  // namespace

extern "C" int psh_cnvt_device(const psh_cnvt *cv, psh_data *const *d_data, const size_t *n,
                               int narr, int *order_out) {
  if (!cv || narr < 0 || (narr && (!d_data || !n))) return PSH_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PSH_ERR_DEVICE;
  hipStream_t s = nullptr;
  const Cosmo c{cv->omega_m, cv->omega_l, cv->omega_k,
                cv->eos_w == -1 ? 0.0 : 3 * (1 + cv->eos_w)};   // src/cnvt_coord.c:498
  if (cv->zsp) {
    if (!cv->dsp || cv->nsp < 2) return PSH_ERR_ARG;
    std::vector<double> ypp = spline_ypp(cv->zsp, cv->dsp, cv->nsp);
    double *tab = nullptr;
    if (hipMalloc(&tab, sizeof(double) * 3 * cv->nsp) != hipSuccess) return PSH_ERR_MEMORY;
    bool ok = hipMemcpy(tab, cv->zsp, sizeof(double) * cv->nsp, hipMemcpyHostToDevice) ==
                  hipSuccess &&
              hipMemcpy(tab + cv->nsp, cv->dsp, sizeof(double) * cv->nsp,
                        hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(tab + 2 * cv->nsp, ypp.data(), sizeof(double) * cv->nsp,
                        hipMemcpyHostToDevice) == hipSuccess;
    for (int a = 0; ok && a < narr; a++) {
      if (!n[a]) continue;
      hipLaunchKernelGGL(interp_kernel, dim3(blocks_for(n[a])), dim3(256), 0, s,
                         (double *)d_data[a], n[a], tab, tab + cv->nsp, tab + 2 * cv->nsp,
                         cv->nsp);
      ok = hipGetLastError() == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(s) == hipSuccess;
    (void)hipFree(tab);
    if (order_out) *order_out = 0;
    return ok ? PSH_OK : PSH_ERR_DEVICE;
  }
  // redshift range of everything to be converted (src/cnvt_coord.c:138-226)
  double zmin = DBL_MAX, zmax = -DBL_MAX;
  double *part = nullptr;
  const int pb = 1024;
  if (hipMalloc(&part, sizeof(double) * PSH_BOUNDS_STRIDE * pb) != hipSuccess)
    return PSH_ERR_MEMORY;
  std::vector<double> h(PSH_BOUNDS_STRIDE * pb);
  for (int a = 0; a < narr; a++) {
    if (!n[a]) continue;
    int blocks = (int)((n[a] + 255) / 256);
    if (blocks > pb) blocks = pb;
    if (psh_coord_bounds((const double *)d_data[a], n[a], part, blocks, s) != hipSuccess ||
        hipMemcpy(h.data(), part, sizeof(double) * PSH_BOUNDS_STRIDE * blocks,
                  hipMemcpyDeviceToHost) !=
            hipSuccess) {
      (void)hipFree(part);
      return PSH_ERR_DEVICE;
    }
    for (int b = 0; b < blocks; b++) {
      zmin = fmin(zmin, h[PSH_BOUNDS_STRIDE * b + 2]);
      zmax = fmax(zmax, h[PSH_BOUNDS_STRIDE * b + 5]);
    }
  }
  (void)hipFree(part);
  if (zmin > zmax) {   // nothing to convert
    if (order_out) *order_out = 0;
    return PSH_OK;
  }
  if (zmin < 0) return PSH_ERR_CNVT;   // negative redshift
  const int order = choose_order(c, cv->ecdst, zmin, zmax);
  if (order < 0) return PSH_ERR_CNVT;
  const RuleDev r = rule_dev(order);
  for (int a = 0; a < narr; a++) {
    if (!n[a]) continue;
    hipLaunchKernelGGL(integr_kernel, dim3(blocks_for(n[a])), dim3(256), 0, s,
                       (double *)d_data[a], n[a], c, r);
    if (hipGetLastError() != hipSuccess) return PSH_ERR_DEVICE;
  }
  if (hipStreamSynchronize(s) != hipSuccess) return PSH_ERR_DEVICE;
  if (order_out) *order_out = order;
  return PSH_OK;
}

extern "C" int psh_cnvt_coord(const psh_cnvt *cv, psh_cata *cat) {
  if (!cv || !cat) return PSH_ERR_ARG;
  std::vector<psh_data *> host, dev;
  std::vector<size_t> cnt;
  for (int i = 0; i < cat->num && i < 2; i++) {
    if (cv->dcnvt[i] && cat->ndata[i]) {
      host.push_back(cat->data[i]);
      cnt.push_back(cat->ndata[i]);
    }
    if (cv->rcnvt[i] && cat->rand && cat->nrand && cat->nrand[i]) {
      host.push_back(cat->rand[i]);
      cnt.push_back(cat->nrand[i]);
    }
  }
  int rc = PSH_OK;
  for (size_t a = 0; a < host.size(); a++) {
    psh_data *d = nullptr;
    if (hipMalloc(&d, sizeof(psh_data) * cnt[a]) != hipSuccess ||
        hipMemcpy(d, host[a], sizeof(psh_data) * cnt[a], hipMemcpyHostToDevice) !=
            hipSuccess) {
      if (d) (void)hipFree(d);
      rc = PSH_ERR_DEVICE;
      break;
    }
    dev.push_back(d);
  }
  if (rc == PSH_OK) rc = psh_cnvt_device(cv, dev.data(), cnt.data(), (int)dev.size(), nullptr);
  for (size_t a = 0; a < dev.size(); a++) {
    if (rc == PSH_OK &&
        hipMemcpy(host[a], dev[a], sizeof(psh_data) * cnt[a], hipMemcpyDeviceToHost) !=
            hipSuccess)
      rc = PSH_ERR_DEVICE;
    (void)hipFree(dev[a]);
  }
  return rc;
}
