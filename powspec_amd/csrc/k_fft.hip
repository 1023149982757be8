// k_fft.hip -- 3-D real-to-complex FFT of the density meshes on gfx950,
// replacing the reference's FFTW r2c/c2r plans (src/fftw_define.h:47-64,
// executed at src/multipole.c:444,459,493 and src/mp_template.c:98), with the
// per-mode work of the reference fused into the last (x) pass:
//   * interlace combine (src/multipole.c:461-487),
//   * |k| / mu binning with Legendre weights for boxes
//     (src/multipole.c:778-898, count_mode_sim_lin),
//   * survey binning and the Y_lm(k) accumulation of the multipole fields
//     (src/multipole.c:579-660, src/mp_template.c:104-137).
//
// Passes over a mesh (layout in psh_internal.h):
//   z: in-place r2c of each row (N/2-point complex FFT + split post-process)
//   y: strided column FFTs over 16-column kz tiles
//   x: strided column FFTs fused with the epilogue; the transformed field is
//      never written back for boxes.
// Unnormalised transforms with FFTW's sign convention (-1 forward).
#include "fft_core.h"
#include "psh_internal.h"

#include <math.h>
#include <type_traits>
#include <vector>

using namespace pshfft;

namespace {

// Schedules of the column pass (y), chosen per workload by the tuner
// (psh_autotune_fft): columns per tile, values per thread and radix plan.
//   0  16 columns (one 128-byte line per row), 32 values, plan 0 (default)
//   1  16 columns, 32 values, plan 1 (smaller radices first)
//   2  16 columns, 32 values, plan 2 (one pass more)
//   3  16 columns, 16 values (twice the threads), plan 0
//   4  32 columns (two lines per row), 32 values, plan 0
//   5  8 columns (half lines), 16 values, plan 0
// 8-column tiles read at about 60 % of the bandwidth of 16-column ones at
// N = 1024; the other shapes trade LDS, registers and waves per CU.
constexpr int kColScheds = 6;
template <int N, int S = 0>
struct ColCfg {
  static constexpr int C = S == 4 ? 32 : (S == 5 ? 8 : 16);
  static constexpr int VPT = (S == 3 || S == 5) ? 16 : 32;
  static constexpr int PL = (S == 1 || S == 2) ? S : 0;
  static constexpr int T0 = C * N / VPT;
  static constexpr int T = T0 < 64 ? 64 : (T0 > 1024 ? 1024 : pow2_ceil(T0));
  static constexpr size_t LDS = sizeof(float) * Lds<N, C, true>::SIZE;
  static constexpr bool OK = LDS <= 160 * 1024 && (S != 2 || npasses(N, 2) > npasses(N, 0)) &&
                             (S != 1 || radix(N, 0, 1) != radix(N, 0, 0));
};

template <int N>
struct ZCfg {
  static constexpr int M = N / 2;
  static constexpr int C = (4096 / M) > 0 ? pow2_floor(4096 / M) : 1;
  static constexpr int T = 256;
};

// ---------------------------------------------------------------------------
// Real-space Y_lm weights of the survey estimator (src/mp_template.c:69-95),
// evaluated in FP32 from per-axis position tables.
// ---------------------------------------------------------------------------
// A real-form Y_lm as norm * sin^|m|(theta) * Q(cos theta) * trig(|m| phi),
// Q the polynomial of degree l - |m| that the associated Legendre recursion
// below produces divided by sin^|m| (coefficients built on the host, norm
// folded in), trig = cos for m > 0, sin for m < 0, 1 for m = 0.  Horner in
// cos theta replaces the recursion's per-point divisions.
struct YlmPoly {
  int am;        // |m|
  int deg;       // l - |m|
  int sinpart;   // m < 0
  float q[8];    // coefficients of Q, ascending powers, times norm
};

struct YlmDev {
  int ell;                // 0: no weighting
  float org[3], step[3];  // r-space position of cell i: org + i step per axis
  YlmPoly poly[1];        // the (l, m) weight, norm folded in
};

// Host: Q from the associated Legendre recursion without the Condon-Shortley
// phase (P_m^m = (2m-1)!! sin^m, P_{m+1}^m = (2m+1) cos P_m^m, P_l^m =
// ((2l-1) cos P_{l-1}^m - (l+m-1) P_{l-2}^m) / (l-m)), run on polynomials in
// cos theta in FP64 with the common sin^m factored out.
inline YlmPoly make_ylm_poly(int ell, int m, float norm) {
  YlmPoly p{};
  const int am = m < 0 ? -m : m;
  p.am = am;
  p.deg = ell - am;
  p.sinpart = m < 0;
  double pmm[8] = {0}, pm1[8] = {0}, pl[8];
  double df = 1.0;
  for (int i = 1; i <= am; i++) df *= 2 * i - 1;
  pmm[0] = df;
  const double *res = pmm;
  if (ell > am) {
    for (int k = 0; k < 7; k++) pm1[k + 1] = (2 * am + 1) * pmm[k];
    for (int l = am + 2; l <= ell; l++) {
      for (int k = 0; k < 8; k++) pl[k] = -(double)(l + am - 1) * pmm[k];
      for (int k = 0; k < 7; k++) pl[k + 1] += (2 * l - 1) * pm1[k];
      for (int k = 0; k < 8; k++) {
        pl[k] /= (double)(l - am);
        pmm[k] = pm1[k];
        pm1[k] = pl[k];
      }
    }
    res = pm1;
  }
  for (int k = 0; k < 8; k++) p.q[k] = (float)(res[k] * norm);
  return p;
}

// Y_lm at the unit vector (ux, uy, uz): sin^|m| theta e^{i |m| phi} is
// (ux + i uy)^|m|, so neither sin theta nor the azimuth is formed.  Horner
// runs over all seven coefficients (those above the degree are 0) and the
// power is a loop of |m| steps, uniform over the launch: selects over the
// degree and |m| per point cost more than the arithmetic they skip.
__device__ __forceinline__ float ylm_poly_eval(const YlmPoly &p, float uz, float ux,
                                               float uy) {
  float q = p.q[6];
#pragma unroll
  for (int k = 5; k >= 0; k--) q = __builtin_fmaf(q, uz, p.q[k]);
  float c = 1.f, s = 0.f;
#pragma unroll 1
  for (int i = 0; i < p.am; i++) {
    const float c2 = c * ux - s * uy;
    s = s * ux + c * uy;
    c = c2;
  }
  return q * (p.sinpart ? s : c);
}

__device__ __forceinline__ float ylm_r_weight(const YlmDev &y, const YlmPoly &p, int i,
                                              int j, int k) {
  if (i == 0 && j == 0) return 1.f;  // first row unweighted (mp_template.c:75)
  // positions from their affine form, not from tables: no load (and no
  // wait for one) per point
  const float rx = __builtin_fmaf((float)i, y.step[0], y.org[0]);
  const float ry = __builtin_fmaf((float)j, y.step[1], y.org[1]);
  const float rz = __builtin_fmaf((float)k, y.step[2], y.org[2]);
  // r^2 is far from the denormal range: the bare reciprocal square root
  const float ir = __builtin_amdgcn_rsqf(rx * rx + ry * ry + rz * rz);
  return ylm_poly_eval(p, rz * ir, rx * ir, ry * ir);
}

// ---------------------------------------------------------------------------
// z pass: real rows -> half-complex rows
// ---------------------------------------------------------------------------
template <int N, bool YLM, bool NT>
__global__ void __launch_bounds__(256)
z_r2c_kernel(const float *__restrict__ in, float2 *__restrict__ out,
             size_t nlines, size_t mstride_f, int H,
             const float2 *__restrict__ tw, YlmDev ylm, int kzlim) {
  constexpr int M = N / 2, C = ZCfg<N>::C, T = ZCfg<N>::T;
  constexpr int V = max_values<M, C, T, false>();
  __shared__ __align__(16) float lds[Lds<M, C, false>::SIZE];
  const size_t line0 = (size_t)blockIdx.x * C;
  const size_t lpm = (size_t)N * N;
  float2 v[V];

  // Line bases.  When every butterfly of the first pass of a line falls in
  // one wave (NBF a multiple of 64: N >= 1024), a wave's line is uniform:
  // its row base is formed once in scalar registers and each element is a
  // 32-bit byte offset from it (SGPR base + VGPR offset loads), instead of
  // 64-bit address arithmetic per element.
  constexpr bool UNIFORM = PassGeom<M, C, T, 0, false>::NBF % 64 == 0 && T % 64 == 0;
  auto row_of = [&](size_t line, size_t &rowi) -> const float2 * {
    const size_t mesh = line / lpm;
    rowi = line - mesh * lpm;
    return reinterpret_cast<const float2 *>(in + mesh * mstride_f + rowi * (size_t)(2 * H));
  };
  // every value of the thread is loaded before any is weighted: a weight
  // computed between a load and its use keeps the compiler from issuing
  // the next loads, and each load's latency is then waited for alone
  pass_load<M, C, T, 0, false>(v, [&](int c, int n) -> float2 {
    const int cu = UNIFORM ? __builtin_amdgcn_readfirstlane(c) : c;
    const size_t line = line0 + cu;
    if (line >= nlines) return make_float2(0.f, 0.f);
    size_t rowi;
    const float2 *row = row_of(line, rowi);
    return UNIFORM ? mesh_load<NT>(row_elem<true>(row, (uint32_t)n)) : mesh_load<NT>(row + n);
  });
  if (YLM) {
    using G0 = PassGeom<M, C, T, 0, false>;
#pragma unroll
    for (int q = 0; q < G0::NB; q++) {
      int c, j0;
      const bool ok = G0::map(q, c, j0);
      const size_t line = line0 + c;
      const size_t rowi = line % lpm;
      const int i = (int)(rowi / N), j = (int)(rowi - (size_t)i * N);
#pragma unroll
      for (int r = 0; r < G0::R; r++) {
        const int n = j0 + r * G0::NBF;
        if (ok && line < nlines) {
          v[q * G0::R + r].x *= ylm_r_weight(ylm, ylm.poly[0], i, j, 2 * n);
          v[q * G0::R + r].y *= ylm_r_weight(ylm, ylm.poly[0], i, j, 2 * n + 1);
        }
      }
    }
  }
  Passes<M, C, T, 0, false, false, 2>::run(v, lds, tw);

  // post-process: X[k] = E + W^k O, X[M-k] = conj(E - W^k O)
  constexpr int NP = M / 2 + 1;
  constexpr int NI = (C * NP + T - 1) / T;
  // line by line (N >= 1024): k = thread + u T over [0, M/2) for one line at
  // a time, so a line's output row is a uniform base again; X[M/2] of each
  // line is formed by one thread at the end
  constexpr bool LINEWISE = UNIFORM && (M / 2) % T == 0;
  constexpr int KPT = LINEWISE ? (M / 2) / T : 1;
  constexpr int NR = LINEWISE ? C * KPT : NI;
  float rk[NR], rmk[NR];
  float rmid = 0.f;
  store_last<M, C, T, false>(v, [&](int c, int n, float2 x) {
    lds[Lds<M, C, false>::idx(c, n)] = x.x;
  });
  __syncthreads();
  if constexpr (LINEWISE) {
    const int t = opaque_tid<T>();
#pragma unroll
    for (int c = 0; c < C; c++)
#pragma unroll
      for (int u = 0; u < KPT; u++) {
        const int k = t + u * T;
        rk[c * KPT + u] = lds[Lds<M, C, false>::idx(c, k)];
        rmk[c * KPT + u] = lds[Lds<M, C, false>::idx(c, mirror<M>(k))];
      }
    if (t < C) rmid = lds[Lds<M, C, false>::idx(t, M / 2)];
  } else {
#pragma unroll
    for (int it = 0; it < NI; it++) {
      int item = threadIdx.x + it * T;
      int c = item / NP, k = item - c * NP;
      if (item < C * NP) {
        rk[it] = lds[Lds<M, C, false>::idx(c, k)];
        rmk[it] = lds[Lds<M, C, false>::idx(c, mirror<M>(k))];
      }
    }
  }
  __syncthreads();
  store_last<M, C, T, false>(v, [&](int c, int n, float2 x) {
    lds[Lds<M, C, false>::idx(c, n)] = x.y;
  });
  __syncthreads();
  if constexpr (LINEWISE) {
    const int t = opaque_tid<T>();
    float2 twk[KPT];
#pragma unroll
    for (int u = 0; u < KPT; u++) twk[u] = tw[t + u * T];
#pragma unroll
    for (int c = 0; c < C; c++) {
      const size_t line = line0 + c;
      if (line >= nlines) break;   // uniform: lines past the end are all at the end
      size_t rowi;
      float2 *orow = out + (line / lpm) * (mstride_f / 2);
      rowi = line - (line / lpm) * lpm;
      orow += rowi * (size_t)H;
#pragma unroll
      for (int u = 0; u < KPT; u++) {
        const int k = t + u * T;
        const float ik = lds[Lds<M, C, false>::idx(c, k)];
        const float imk = lds[Lds<M, C, false>::idx(c, mirror<M>(k))];
        const float r0 = rk[c * KPT + u], r1 = rmk[c * KPT + u];
        const float2 E = make_float2(0.5f * (r0 + r1), 0.5f * (ik - imk));
        const float2 O = make_float2(0.5f * (ik + imk), -0.5f * (r0 - r1));
        const float2 WO = cmul(twk[u], O);
        if (k <= kzlim) mesh_store<NT>(row_elem<true>(orow, (uint32_t)k), cadd(E, WO));
        const float2 d = csub(E, WO);   // k < M/2: M - k != k
        if (M - k <= kzlim)
          mesh_store<NT>(row_elem<true>(orow, (uint32_t)(M - k)), make_float2(d.x, -d.y));
      }
    }
    if (t < C && line0 + t < nlines && M / 2 <= kzlim) {   // X[M/2]: one value per line
      const size_t line = line0 + t;
      const size_t mesh = line / lpm, rowi = line - mesh * lpm;
      float2 *orow = out + mesh * (mstride_f / 2) + rowi * (size_t)H;
      constexpr int k = M / 2;
      const float ik = lds[Lds<M, C, false>::idx(t, k)];
      const float2 E = make_float2(0.5f * (rmid + rmid), 0.5f * (ik - ik));
      const float2 O = make_float2(0.5f * (ik + ik), -0.5f * (rmid - rmid));
      const float2 WO = cmul(tw[k], O);
      mesh_store<NT>(orow + k, cadd(E, WO));
    }
    return;
  }
  // staged rows: measured on MI355X, z pass at N = 512 0.267 -> 0.255 ms per
  // mesh pair; at N = 1024 the extra registers (70 -> 86 VGPRs, 7 -> 5 waves
  // per SIMD) cost more than the whole lines save (4.17 -> 4.33 ms)
  constexpr bool STAGE = NT && N == 512;
  if constexpr (!STAGE) {
#pragma unroll
    for (int it = 0; it < NI; it++) {
      int item = threadIdx.x + it * T;
      int c = item / NP, k = item - c * NP;
      if (item >= C * NP) continue;
      size_t line = line0 + c;
      if (line >= nlines) continue;
      float ik = lds[Lds<M, C, false>::idx(c, k)];
      float imk = lds[Lds<M, C, false>::idx(c, mirror<M>(k))];
      float2 E = make_float2(0.5f * (rk[it] + rmk[it]), 0.5f * (ik - imk));
      float2 O = make_float2(0.5f * (ik + imk), -0.5f * (rk[it] - rmk[it]));
      float2 WO = cmul(tw[k], O);
      size_t mesh = line / lpm, rowi = line - mesh * lpm;
      float2 *orow = out + mesh * (mstride_f / 2) + rowi * (size_t)H;
      if (k <= kzlim) mesh_store<NT>(orow + k, cadd(E, WO));
      if (M - k != k && M - k <= kzlim) {
        float2 d = csub(E, WO);
        mesh_store<NT>(orow + (M - k), make_float2(d.x, -d.y));
      }
    }
  } else {
    // The pairs X[k], X[M-k] land in two places of a row, so stores straight
    // from the pairs leave partial 128-byte lines, which non-temporal stores
    // send to HBM unmerged.  The rows are staged in LDS instead (half the
    // block's rows at a time, zeros in the padding) and written as whole
    // lines of 16-byte stores.
    constexpr int HH = M + PSH_PAD;      // complex values per stored row
    constexpr int HALF = C / 2;
    static_assert(C % 2 == 0 && HALF * HH * 2 <= Lds<M, C, false>::SIZE,
                  "half of the block's rows fit the exchange tile");
    static_assert(HH % 2 == 0, "rows of whole float4 pairs");
    float ik[NI], imk[NI];
#pragma unroll
    for (int it = 0; it < NI; it++) {
      const int item = threadIdx.x + it * T;
      const int c = item / NP, k = item - c * NP;
      if (item < C * NP) {
        ik[it] = lds[Lds<M, C, false>::idx(c, k)];
        imk[it] = lds[Lds<M, C, false>::idx(c, mirror<M>(k))];
      }
    }
    float2 *stage = reinterpret_cast<float2 *>(lds);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      __syncthreads();   // the tile's previous contents are consumed
#pragma unroll
      for (int it = 0; it < NI; it++) {
        const int item = threadIdx.x + it * T;
        const int c = item / NP, k = item - c * NP;
        if (item >= C * NP || c / HALF != h) continue;
        const float2 E = make_float2(0.5f * (rk[it] + rmk[it]), 0.5f * (ik[it] - imk[it]));
        const float2 O = make_float2(0.5f * (ik[it] + imk[it]), -0.5f * (rk[it] - rmk[it]));
        const float2 WO = cmul(tw[k], O);
        float2 *srow = stage + (c - h * HALF) * HH;
        srow[k] = cadd(E, WO);
        if (M - k != k) {
          const float2 d = csub(E, WO);
          srow[M - k] = make_float2(d.x, -d.y);
        }
        if (k >= 1 && k < PSH_PAD) srow[M + k] = make_float2(0.f, 0.f);
      }
      __syncthreads();
      constexpr int Q = HH / 2;           // float4 (two complex values) per row
      const float4 *src = reinterpret_cast<const float4 *>(stage);
      for (int i = threadIdx.x; i < HALF * Q; i += T) {
        const int r = i / Q, q = i - r * Q;
        const size_t line = line0 + h * HALF + r;
        if (line >= nlines) continue;
        const size_t mesh = line / lpm, rowi = line - mesh * lpm;
        float4 *orow = reinterpret_cast<float4 *>(out + mesh * (mstride_f / 2) + rowi * (size_t)H);
        if (2 * q <= kzlim) mesh_store4(orow + q, src[i]);
      }
    }
  }
}

// inverse z pass: half-complex rows -> real rows (times scale), in place;
// the imaginary parts of the kz = 0 and kz = N/2 entries are ignored.
template <int N>
__global__ void __launch_bounds__(256)
z_c2r_kernel(float2 *__restrict__ data, size_t nlines, size_t mstride_f, int H,
             const float2 *__restrict__ tw, float scale) {
  constexpr int M = N / 2, C = ZCfg<N>::C, T = ZCfg<N>::T;
  constexpr int V = max_values<M, C, T, false>();
  __shared__ float lds[Lds<M, C, false>::SIZE];
  const size_t line0 = (size_t)blockIdx.x * C;
  const size_t lpm = (size_t)N * N;
  constexpr int NP = M / 2 + 1;
  constexpr int NI = (C * NP + T - 1) / T;
  float2 zk[NI], zmk[NI];
#pragma unroll
  for (int it = 0; it < NI; it++) {
    int item = threadIdx.x + it * T;
    int c = item / NP, k = item - c * NP;
    zk[it] = zmk[it] = make_float2(0.f, 0.f);
    size_t line = line0 + c;
    if (item >= C * NP || line >= nlines) continue;
    size_t mesh = line / lpm, rowi = line - mesh * lpm;
    const float2 *row = data + mesh * (mstride_f / 2) + rowi * (size_t)H;
    float2 xk = row[k], xmk = row[M - k];
    if (k == 0) {
      xk.y = 0.f;
      xmk.y = 0.f;
    }
    // Z'[k] = (X[k] + conj X[M-k]) + i (X[k] - conj X[M-k]) conj(W^k)
    float2 w = tw[k];
    float2 s = make_float2(xk.x + xmk.x, xk.y - xmk.y);
    float2 d = cmulc(make_float2(xk.x - xmk.x, xk.y + xmk.y), w);
    zk[it] = make_float2(s.x - d.y, s.y + d.x);
    // Z'[M-k] = (X[M-k] + conj X[k]) - i (X[M-k] - conj X[k]) W^k
    float2 s2 = make_float2(xmk.x + xk.x, xmk.y - xk.y);
    float2 d2 = cmul(make_float2(xmk.x - xk.x, xmk.y + xk.y), w);
    zmk[it] = make_float2(s2.x + d2.y, s2.y - d2.x);
  }
  __syncthreads();   // all rows read before any LDS use
  float2 v[V];
  // two-phase LDS fill (re, then im), reading pass-0 inputs in between
  float re[V];
#pragma unroll
  for (int phase = 0; phase < 2; phase++) {
#pragma unroll
    for (int it = 0; it < NI; it++) {
      int item = threadIdx.x + it * T;
      int c = item / NP, k = item - c * NP;
      if (item >= C * NP) continue;
      lds[Lds<M, C, false>::idx(c, k)] = phase ? zk[it].y : zk[it].x;
      if (k != 0 && M - k != k) lds[Lds<M, C, false>::idx(c, M - k)] = phase ? zmk[it].y : zmk[it].x;
    }
    __syncthreads();
    pass_load<M, C, T, 0, false>(v, [&](int c, int n) -> float2 {
      return make_float2(lds[Lds<M, C, false>::idx(c, n)], 0.f);
    });
    if (phase == 0) {
#pragma unroll
      for (int q = 0; q < V; q++) re[q] = v[q].x;
    } else {
#pragma unroll
      for (int q = 0; q < V; q++) v[q] = make_float2(re[q], v[q].x);
    }
    __syncthreads();
  }
  Passes<M, C, T, 0, false, true, 2>::run(v, lds, tw);
  store_last<M, C, T, false>(v, [&](int c, int n, float2 x) {
    size_t line = line0 + c;
    if (line >= nlines) return;
    size_t mesh = line / lpm, rowi = line - mesh * lpm;
    float2 *row = data + mesh * (mstride_f / 2) + rowi * (size_t)H;
    row[n] = make_float2(x.x * scale, x.y * scale);
  });
}

// ---------------------------------------------------------------------------
// strided column passes (y: OS = N*H, ES = H; x: OS = H, ES = N*H)
// ---------------------------------------------------------------------------
template <int N, bool INV, bool NT, int S>
__global__ void __launch_bounds__((ColCfg<N, S>::T))
col_kernel(const float2 *__restrict__ in, float2 *__restrict__ out, int ntz,
           size_t OS, size_t ES, size_t mstride_c, const float2 *__restrict__ tw,
           const int *__restrict__ ylive, int cxs) {
  using CC = ColCfg<N, S>;
  constexpr int C = CC::C, T = CC::T, M = N / 2, PL = CC::PL;
  constexpr int V = max_values<N, C, T, true, 0, PL>();
  extern __shared__ float lds[];
  const size_t tile = blockIdx.x;
  const size_t per_mesh = (size_t)N * ntz;
  const size_t mesh = tile / per_mesh;
  const size_t rem = tile - mesh * per_mesh;
  const int o = (int)(rem / ntz), q = (int)(rem - (size_t)o * ntz);
  const size_t base = mesh * mstride_c + o * OS + (size_t)q * C;
  const int ncv = min(C, M + 1 - q * C);
  if (ylive) {
    // a tile of k_z columns with no live k_y at all (k_z past k_max) is
    // never read by the x pass: nothing to transform
    bool any = false;
    for (int t = (q * C) >> cxs; t <= (q * C + ncv - 1) >> cxs; t++) any |= ylive[t] >= 0;
    if (!any) return;
  }
  using RS = RowSplit<N, C, T, true, PL>;
  static_assert(PSH_PAD == 16 && RS::XSPAN < (1ull << 32), "lane offsets are 32-bit");
  // columns q*C + c < H always lie inside the padded row (the row holds
  // M + 1 + 15 columns, so a tile of C columns starting at or before M ends
  // inside it when C <= 16; 32-column tiles are clipped below); those past
  // kz = M are read and discarded
  float2 v[V];
  auto ld = [&](int r, int c, int j) -> float2 {
    if (C > 16 && q * C + c >= M + 1 + 15) return make_float2(0.f, 0.f);
    const float2 x = mesh_load<NT>(row_elem<RS::BYTE32>(in + base + (size_t)(r * RS::LOAD) * ES,
                                                         (uint32_t)j * (uint32_t)ES + c));
    return c < ncv ? x : make_float2(0.f, 0.f);
  };
  transform_rows<N, C, T, true, INV, 1, decltype(ld), PL>(v, lds, tw, ld);
  // outputs k_y = i' (i = j + r STORE) past the column's live bound are
  // never read (the fused x pass skips their tiles): not stored
  int kylim = M;
  if (ylive) {
    const int cc = min(opaque_tid<T>() % C, ncv - 1);   // the thread's column in every pass
    kylim = ylive[(q * C + cc) >> cxs];
  }
  auto st = [&](int r, int c, int j, float2 x) {
    const int i = j + r * RS::STORE;
    const int ay = i <= M ? i : N - i;
    if (c < ncv && ay <= kylim)
      mesh_store<NT>(row_elem<RS::BYTE32>(out + base + (size_t)(r * RS::STORE) * ES,
                                          (uint32_t)j * (uint32_t)ES + c), x);
  };
  store_last_rows<N, C, T, true, decltype(st), PL>(v, st);
}

// ---------------------------------------------------------------------------
// x pass fused with interlacing and binning
// ---------------------------------------------------------------------------
struct XBinDev {
  const float2 *A, *B, *F0;
  float2 *store;
  size_t mstride_c;
  PshAxisTables ax;
  PshBins bins;
  const float2 *phase;
  double *partials;
  int nacc;        // accumulators of the launch's bin window (LDS)
  int nacc_full;   // accumulators per workgroup row of partials (all bins)
  int kb0, nbw;    // bin window [kb0, kb0 + nbw) of the launch
  double xbound;   // (largest |pair contribution| / phase 1's max) x N C, with margins
  int losx0;       // the line of sight has no x component
  int has_dc;      // replace the k = 0 mode by dcf
  float dcf;
  // SURVEYL: nf fields of one pole l (at A + f fstride_c), field f weighted
  // by the real-form Y_lm(k) of yp[f]
  int nf;
  size_t fstride_c;
  YlmPoly yp[PSH_MAX_YLM_FIELDS];
  float vec0f;   // 2 pi / L_x
  float kd0f;    // 1 / L_x (Y_lm(k) direction; k_y, k_z from the FP64 tables)
  // SIM: L_ell(mu) = mu^(ell & 1) * (((c3 x + c2) x + c1) x + c0), x = mu^2,
  // for the requested pole l; bit l of oddmask set when ell is odd
  float legc[PSH_MAX_NL][4];
  int oddmask;
  int even024;   // the poles are exactly 0, 2, 4 (the common request)
  const int *ylive;   // per k_z tile, the largest |k_y| with binned modes, or nullptr
};

// Bin of a mode from |k|^2, in the reference's operation order
// (src/multipole.c:147-158); -1 when outside the bins.
__device__ __forceinline__ int mode_bin(double kmod2, const PshBins &b,
                                        double &kmod) {
#pragma clang fp contract(off)
  kmod = b.logscale ? 0.5 * log10(kmod2) : sqrt(kmod2);
  if (kmod < b.kedge0 || kmod >= b.kedgeN) return -1;
  int kb = (int)((kmod - b.kedge0) / b.dk);
  if (kb < 0 || kb >= b.nbin) return -1;
  return kb;
}

// x pass: up to 1024 threads (16 waves, 128 VGPRs).  The plain kernel takes
// 16-column tiles (128-byte row segments) and 16 values per thread.  The
// interlaced kernel transforms both fields pass by pass together, also in
// 16-column tiles up to N = 1024 (16 values per thread and field there; the
// register limit spills a few values outside the butterflies, which costs
// far less than the 50 % of extra HBM reads of 8-column tiles: 5.05 -> 4.25
// ms at N = 1024); at N = 2048 8 columns (64-byte half lines, completed by
// the neighbouring tile on the same XCD, see xcd_tile).  With T a multiple of
// C, a thread works on the same column (kz) in every pass and in the binning
// loop.
template <int N, bool IL>
struct XCfg {
  static constexpr int C = N <= 1024 ? 16 : 8;
  // up to N = 256 the first pass's radix-16 butterflies set the values per
  // thread whatever T is: T = its butterfly count leaves no thread idle
  static constexpr bool SMALL = IL && N <= 256;
  static constexpr int T0 = SMALL ? C * N / radix(N, 0) : C * N / (IL ? 8 : 16);
  // a power of two: grid sizes 3 2^k and 5 2^k round up, which leaves some
  // threads idle in the power-of-two passes but keeps the values per
  // thread (and the registers) of the power-of-two grids
  static constexpr int T = T0 < 64 ? 64 : (T0 > 1024 ? 1024 : pow2_ceil(T0));
  // waves per SIMD the register allocation must allow: 4 (128 VGPRs) lets
  // workgroups share a CU and hide each other's barriers; the small
  // interlaced kernel holds 2 x 16 values per thread and gets 2 (256 VGPRs)
  static constexpr int WAVES = SMALL ? 2 : 4;
  // the next tile's rows loaded during phase 2 (see x_bin_kernel): 2 every
  // field, 1 the first field only (the second loads when the tile starts),
  // 0 none.  The interlaced kernel from N = 640 holds both fields' 16 values
  // per thread and spills registers when both prefetches are in flight as
  // well (measured on MI355X at N = 1024: 4.31 ms with both, 4.00 ms with
  // the first field's, 3.46 ms without)
  static constexpr int PREFETCH = (IL && N >= 640) ? 0 : 2;
  static_assert(T % C == 0, "a thread keeps one column");
};

// LDS of the x pass besides the bins: the tile maximum, ksq_x (FP64; read
// from global memory by the GS layout), xtab, twiddles, exchange tile.  The twiddles are read from LDS: read from global memory
// they are invariant over the persistent tile loop and stay live in
// registers.
template <int N, bool IL, bool GS = false>
constexpr size_t xbin_static_lds() {
  return sizeof(double) * 2 + (GS ? 0 : sizeof(double) * N) + sizeof(float4) * N +
         sizeof(float2) * N + sizeof(float) * 4 * PSH_MAX_NL +
         sizeof(float) * Lds<N, XCfg<N, IL>::C, true>::SIZE;
}

// LDS per accumulator of the x pass: the workgroup's FP64 sum and the
// current tile's 64-bit integer sum (kXbinAccBytesG when the FP64 sums are
// kept in the workgroup's row of the partials instead, the GS layout).
constexpr size_t kXbinAccBytes = sizeof(double) + sizeof(long long);
constexpr size_t kXbinAccBytesG = sizeof(long long);
constexpr size_t kXbinLdsMax = 160 * 1024;

// Adds x * 2^-E, rounded to the nearest integer, to a 64-bit integer sum in
// LDS.  sdown = 2^-E; |x| sdown < 2^51, so adding 1.5 * 2^52 leaves the
// rounded integer in the low mantissa bits of the double (exact scaling,
// one rounding) and the subtraction of the constant's bits recovers it.
// Integer additions are associative: the tile's sums do not depend on the
// order in which the lanes and waves arrive.
__device__ __forceinline__ void acc_fixed(long long *ib, float x, double sdown) {
  const double y = __builtin_fma((double)x, sdown, 6755399441055744.0);
  const long long q = __double_as_longlong(y) - 0x4338000000000000LL;
  atomicAdd(reinterpret_cast<unsigned long long *>(ib), (unsigned long long)q);
}

// Bin of a mode from its FP32 |k|, kmodf: -1 outside the bins, -2 when
// undecided.  t = (|k| - kedge0) / dk evaluated in FP32 errs by less than
// 2e-6 (|t| + kedge0/dk + 8), so when t lies farther than that from an
// integer, floor(t) is the bin the FP64 expression gives.  Log bins are
// always undecided.
__device__ __forceinline__ int bin_fast(const PshBins &b, float kmodf) {
  const float tf = (kmodf - b.kedge0f) * b.invdkf;
  const float fl = floorf(tf);
  const float fr = tf - fl;
  const float eps = 2e-6f * (fabsf(tf) + b.k0dk + 8.f);
  if (b.logscale || !(fr > eps && fr < 1.f - eps)) return -2;
  return (fl >= 0.f && fl < (float)b.nbin) ? (int)fl : -1;
}

// The FP64 bin of the reference: |k|^2 = (ksq_x + ksq_y) + ksq_z with ksq_z
// added only off the kz = 0 plane (src/multipole.c:147-158).
__device__ __forceinline__ int bin_exact(const PshBins &b, double ksqx, double ksqy,
                                      double ksqz, int kz) {
#pragma clang fp contract(off)
  double kmod2 = ksqx + ksqy;
  if (kz > 0) kmod2 = kmod2 + ksqz;
  double kmod;
  return mode_bin(kmod2, b, kmod);
}

// Forward x transform of one (y, kz-tile) column tile fused with the
// interlace combine and the binning of its modes.  Two phases per tile:
//  1. (registers, unrolled) F = (F_A + e^{i pi (i'+j'+k)/N} F_B) / 2
//     (src/multipole.c:476-484); each mode reduced to one float -- |F|^2, or
//     Re(F0 conj F) for l > 0 survey fields -- written to LDS in natural
//     order, F itself stored for the l = 0 survey field;
//  2. (LDS, one loop over the pairs of modes k_x = +-i of the column, which
//     share |k|, the bin and, for a line of sight without an x component,
//     mu; only the alias factor differs, src/multipole.c:50-58) bin, alias
//     correction, half-space multiplicity and Legendre weights
//     (src/multipole.c:147-177, src/mp_template.c:107-139), accumulated as
//     64-bit integers at a power-of-two scale set by the tile's largest
//     |v|, so that every sum of the tile fits; after the tile each
//     accumulator's integer is converted and added to its FP64 sum by the
//     thread that owns it.  The sums are therefore the same on every run.
// xtab[i] = {alias_x, ksq_x, phase_x.re, phase_x.im} in LDS.
// GS: the workgroup's FP64 sums live in its row of the partials and ksq_x
// is read from global memory, which halves the LDS (two workgroups per CU
// for the plain pass with |mu| cells).
template <int N, int EPI, bool INTLACE, bool NT, bool GS>
__global__ void __launch_bounds__((XCfg<N, INTLACE>::T), (XCfg<N, INTLACE>::WAVES))
x_bin_kernel(XBinDev a, int ntz, int H, const float2 *__restrict__ tw) {
  constexpr int C = XCfg<N, INTLACE>::C, T = XCfg<N, INTLACE>::T, M = N / 2;
  constexpr int V = max_values<N, C, T, true>();
  using L = Lds<N, C, true>;
  using G = PassGeom<N, C, T, npasses(N) - 1, true>;
  extern __shared__ double dyn[];
  const int nacc_pad = (a.nacc + 1) & ~1;
  long long *ibins = reinterpret_cast<long long *>(dyn);
  double *bins = GS ? nullptr : dyn + nacc_pad;
  unsigned *tmaxp = reinterpret_cast<unsigned *>(dyn + (GS ? 1 : 2) * nacc_pad);
  double *ksqx = dyn + (GS ? 1 : 2) * nacc_pad + 2;   // unused by GS
  float4 *xtab = reinterpret_cast<float4 *>(ksqx + (GS ? 0 : N));
  float2 *twl = reinterpret_cast<float2 *>(xtab + N);
  // the Legendre coefficients of the general pole loop, read from LDS: held
  // in scalar registers they spill most of the kernel's other scalars
  float4 *legl = reinterpret_cast<float4 *>(twl + N);
  float *lds = reinterpret_cast<float *>(legl + PSH_MAX_NL);
  const PshBins &b = a.bins;
  const int mock = blockIdx.y;
  // the workgroup's row of the partials; window accumulator i (pole slot l,
  // bin kb0 + rest) is its entry (l nbin + kb0) nmu + rest
  double *out = a.partials + ((size_t)mock * gridDim.x + blockIdx.x) * a.nacc_full;
  const int nmu_o = EPI == PSH_EPI_SIM ? b.nmu : 1;
  const int stride_o = a.nbw * nmu_o;
  auto out_idx = [&](int i) -> size_t {
    const int l = i / stride_o;
    return (size_t)(l * b.nbin + a.kb0) * nmu_o + (i - l * stride_o);
  };
  for (int i = threadIdx.x; i < a.nacc; i += T) {
    ibins[i] = 0;
    if (GS) out[out_idx(i)] = 0.0;
    else bins[i] = 0.0;
  }
  if (threadIdx.x == 0) *tmaxp = 0u;
  for (int i = threadIdx.x; i < N; i += T) {
    const float2 ph = a.phase[i];
    xtab[i] = make_float4(a.ax.alf[0][i], a.ax.ksqf[0][i], ph.x, ph.y);
    if (!GS) ksqx[i] = a.ax.ksq[0][i];
    twl[i] = tw[i];
  }
  if (EPI == PSH_EPI_SIM && threadIdx.x < PSH_MAX_NL) {
    const int l = threadIdx.x;
    legl[l] = make_float4(a.legc[l][0], a.legc[l][1], a.legc[l][2], a.legc[l][3]);
  }
  __syncthreads();

  const size_t ES = (size_t)N * H;
  const uint32_t ES32 = (uint32_t)ES;
  using RS = RowSplit<N, C, T, true>;
  static_assert(G::NS == RS::STORE, "phase 1 walks the last pass outputs");
  static_assert(PSH_PAD == 16 && RS::XSPAN < (1ull << 32), "lane offsets are 32-bit");
  const size_t ntiles = (size_t)N * ntz;
  const size_t t0 = xcd_tile(blockIdx.x, gridDim.x);
  // tile -> (y, kz block q); rows of the block start at base
  auto tile_base = [&](size_t t, int &y, int &q) -> size_t {
    y = (int)(t / ntz);
    q = (int)(t - (size_t)y * ntz);
    return mock * a.mstride_c + (size_t)y * H + (size_t)q * C;
  };
  // columns q*C + c < H always lie inside the padded row; those past
  // kz = M are read and discarded
  auto ld_field = [&](const float2 *f, size_t base, int ncv) {
    return [=](int r, int cc, int j) -> float2 {
      const float2 x = mesh_load<NT>(
          row_elem<RS::BYTE32>(f + base + (size_t)(r * RS::LOAD) * ES, (uint32_t)j * ES32 + cc));
      return cc < ncv ? x : make_float2(0.f, 0.f);
    };
  };
  // The next tile's rows are loaded into va/vb as soon as phase 1 has
  // consumed them, so their HBM latency hides behind phase 2 (one block per
  // CU: nothing else would overlap it) -- as far as the registers hold them
  // (XCfg::PREFETCH: the fields not prefetched load when the tile starts).
  constexpr int PF = XCfg<N, INTLACE>::PREFETCH;
  float2 va[V];
  float2 vb[INTLACE ? V : 1];
  auto load_field = [&](size_t t, bool second) {
    int ty, tq;
    const size_t tb = tile_base(t, ty, tq);
    const int tn = min(C, M + 1 - tq * C);
    if (!second) pass_load_rows<N, C, T, true>(va, ld_field(a.A, tb, tn));
    else if (INTLACE) pass_load_rows<N, C, T, true>(vb, ld_field(a.B, tb, tn));
  };
  // field f of a SURVEYL group (f = 0 is A)
  auto load_group_field = [&](size_t t, int f) {
    int ty, tq;
    const size_t tb = tile_base(t, ty, tq);
    const int tn = min(C, M + 1 - tq * C);
    pass_load_rows<N, C, T, true>(va, ld_field(a.A + (size_t)f * a.fstride_c, tb, tn));
  };
  auto prefetch = [&](size_t t) {
    if (PF >= 1) load_field(t, false);
    if (PF >= 2) load_field(t, true);
  };
  // A tile whose nearest mode (k_x = 0, lowest k_z) already lies at or past
  // the last bin edge has no mode in the bins: its x transform is skipped.
  // ylive[q] is the largest |k_y| index of a live tile in k_z tile q (made on
  // the host with the binning's FP64 |k|, which only grows with |k_x|, |k_y|
  // and k_z); with k_max = k_Ny about 1 - pi/4 of the (k_y, k_z) tiles are
  // skipped.  nullptr (the survey l = 0 pass, whose stored field is
  // transformed back in full, and log bins): every tile.
  auto live = [&](size_t t) -> bool {
    if (!a.ylive) return true;
    int ty, tq;
    (void)tile_base(t, ty, tq);
    const int ay = ty <= M ? ty : N - ty;
    return ay <= a.ylive[tq];
  };
  auto next_live = [&](size_t t) -> size_t {
    while (t < ntiles && !live(t)) t += gridDim.x;
    return t;
  };
  size_t tile = next_live(t0);
  if (tile < ntiles) prefetch(tile);
  for (; tile < ntiles;) {
    int y, q;
    const size_t base = tile_base(tile, y, q);
    if (PF < 1) load_field(tile, false);
    if (PF < 2) load_field(tile, true);
    const int ncv = min(C, M + 1 - q * C);
    if (INTLACE)
      Passes2<N, C, T, 0, true, false, 1>::run(va, vb, lds, twl);
    else
      Passes<N, C, T, 0, true, false, 1>::run(va, lds, twl);
    const int c = opaque_tid<T>() % C;   // the thread's column in every pass
    const bool colok = c < ncv;
    const int kz = min(q * C + c, M);

    // l > 0 survey fields: G = sum over the group's fields f of
    // Y_lm_f(k) F_f, one x transform per field, each weighted as it leaves
    // the transform; the mode k_x = k_y = 0 is added unweighted
    // (src/mp_template.c:121).  Binning Re(F0 conj G) once per group
    // replaces a binning per field.
    float2 gsum[EPI == PSH_EPI_SURVEYL ? V : 1];
    if (EPI == PSH_EPI_SURVEYL) {
      const float kdy = (float)a.ax.kd[1][y], kdz = (float)a.ax.kd[2][kz];
#pragma unroll 1
      for (int f = 0; f < a.nf; f++) {
        if (f > 0) {
          // every thread has read the previous transform's last exchange
          __syncthreads();
          load_group_field(tile, f);
          Passes<N, C, T, 0, true, false, 1>::run(va, lds, twl);
        }
        const YlmPoly &yp = a.yp[f];
#pragma unroll
        for (int qq = 0; qq < G::NB; qq++) {
          const int bt = opaque_tid<T>() + qq * T;
          const int j = bt / C;
#pragma unroll
          for (int r = 0; r < G::R; r++) {
            const int i = j + r * G::NS;
            const int si = i <= M ? i : i - N;
            const float kdx = (float)si * a.kd0f;
            const float k2 = kdx * kdx + kdy * kdy;
            float yw = 1.f;
            if (k2 != 0.f) {
              const float ik = __builtin_amdgcn_rsqf(k2 + kdz * kdz);
              yw = ylm_poly_eval(yp, kdz * ik, kdx * ik, kdy * ik);
            }
            const float2 F = va[qq * G::R + r];
            float2 &g = gsum[qq * G::R + r];
            if (f == 0) g = make_float2(0.f, 0.f);
            g = make_float2(__builtin_fmaf(yw, F.x, g.x), __builtin_fmaf(yw, F.y, g.y));
          }
        }
      }
    }

    // phase 1
    float2 pyz = make_float2(1.f, 0.f);
    if (INTLACE) pyz = cmul(a.phase[y], a.phase[kz]);
    // The tile's fixed-point scale follows the contributions it sums: the
    // thread's largest |v| (the simulation's DC mode, never binned, left
    // out) times its column's weight w_yz; the largest x alias factor is in
    // xbound.  Scaled by the DC power and the grid's largest alias factors
    // instead, the shot-level modes of a large grid lost 20 bits.
    const bool dccol = EPI == PSH_EPI_SIM && y == 0 && kz == 0;
    float vmax = 0.f;
    // l > 0 survey fields: every F0 value of the thread is loaded before the
    // first is used, so the loads' latencies overlap
    float2 f0v[EPI == PSH_EPI_SURVEYL ? V : 1];
    if (EPI == PSH_EPI_SURVEYL) {
#pragma unroll
      for (int qq = 0; qq < G::NB; qq++) {
        const int bt = opaque_tid<T>() + qq * T;
        const int j = bt / C;
        const bool ok = colok && bt < G::NBT;
#pragma unroll
        for (int r = 0; r < G::R; r++) {
          const size_t rbase = base + (size_t)(r * RS::STORE) * ES;
          const uint32_t lo = (uint32_t)j * ES32 + c;
          f0v[qq * G::R + r] =
              ok ? *row_elem<RS::BYTE32>(a.F0 + rbase, lo) : make_float2(0.f, 0.f);
        }
      }
    }
#pragma unroll
    for (int qq = 0; qq < G::NB; qq++) {
      const int bt = opaque_tid<T>() + qq * T;
      const int j = bt / C;
      const bool ok = colok && bt < G::NBT;
#pragma unroll
      for (int r = 0; r < G::R; r++) {
        if (!ok) continue;
        const int i = j + r * G::NS;
        float2 F = EPI == PSH_EPI_SURVEYL ? gsum[qq * G::R + r] : va[qq * G::R + r];
        if (INTLACE) {
          const float4 xt = xtab[i];
          const float2 t = cmul(cmul(make_float2(xt.z, xt.w), pyz), vb[qq * G::R + r]);
          // the simulation epilogue takes |F_A + t|^2 and applies the
          // (1/2)^2 with the per-column weight in phase 2
          if (EPI == PSH_EPI_SIM) F = make_float2(F.x + t.x, F.y + t.y);
          else F = make_float2(0.5f * (F.x + t.x), 0.5f * (F.y + t.y));
        }
        // the field's k = 0 mode is the sum of its weights (the windows sum
        // to 1): data - alpha randoms cancel there, so the exact FP64 sum
        // replaces the FP32 transform's rounding-dominated value
        if (EPI == PSH_EPI_SURVEY0 && a.has_dc && y == 0 && kz == 0 && i == 0)
          F = make_float2(a.dcf, 0.f);
        // element (i = j + r NS, c): row base uniform, lane offset j ES + c
        const size_t rbase = base + (size_t)(r * RS::STORE) * ES;
        const uint32_t lo = (uint32_t)j * ES32 + c;
        float v;
        if (EPI == PSH_EPI_SURVEYL) {
          const float2 f0 = f0v[qq * G::R + r];
          v = f0.x * F.x + f0.y * F.y;
        } else {
          v = F.x * F.x + F.y * F.y;
        }
        if (EPI == PSH_EPI_SURVEY0 && a.store) *row_elem<RS::BYTE32>(a.store + rbase, lo) = F;
        lds[L::idx(c, j) + L::off(r * G::NS)] = v;   // i = j + r NS
        vmax = fmaxf(vmax, (dccol && i == 0) ? 0.f : fabsf(v));
      }
    }
    // largest |v| w_yz of the tile (non-negative floats order as their bits)
    if (colok) {
      vmax *= a.ax.alf[1][y] * a.ax.alf[2][kz] * ((kz == 0 || kz == M) ? 1.f : 2.f) *
              (INTLACE && EPI == PSH_EPI_SIM ? 0.25f : 1.f);
      if (vmax > 0.f) atomicMax(tmaxp, __float_as_uint(vmax));
    }
    __syncthreads();
    const size_t next = next_live(tile + gridDim.x);
    if (next < ntiles) prefetch(next);

    // phase 2: column c, pairs (i, N - i) with i = threadIdx.x / C + k T / C
    // up to M.  The tile's integer sums take the scale 2^E with
    // tmax xbound < 2^(E + 61): every contribution (a pair's, w + w') is
    // below tmax xbound / (N C) and an accumulator takes at most N C of
    // them, so no sum leaves 2^61 and each contribution stays below 2^51
    // (acc_fixed).
    int E = 0;
    {
      const float tmax = __uint_as_float(*tmaxp);
      if (tmax > 0.f) {
        int e;
        (void)frexp((double)tmax * a.xbound, &e);
        E = e - 61;
      }
    }
    const double sdown = ldexp(1.0, -E);
    if (colok) {
      const float ksqyz = a.ax.ksqf[1][y] + a.ax.ksqf[2][kz];
      const float dyz = a.ax.kvf[1][y] * b.losf[1] + a.ax.kvf[2][kz] * b.losf[2];
      const bool single = (kz == 0) || (kz == M);
      const float wyz = a.ax.alf[1][y] * a.ax.alf[2][kz] *
                        ((single ? 1.f : 2.f) * (INTLACE && EPI == PSH_EPI_SIM ? 0.25f : 1.f));
      const double ksqy = a.ax.ksq[1][y], ksqz = a.ax.ksq[2][kz];
      const bool dcrow = (EPI == PSH_EPI_SIM) && y == 0 && kz == 0;
      // the two half-spaces cancel odd poles off the kz = 0, M planes
      const int lmask = single ? ~0 : ~a.oddmask;
      constexpr int S = T / C;   // rows between a thread's modes
      const int i0 = opaque_tid<T>() / C;
      const float *vrow = lds + L::idx(c, i0);
      // One mode pair: its bin, weights and accumulations.  xt, v, vp, ap
      // come from LDS, read one pair ahead by the loop below.
      auto mode_pair = [&](int i, float4 xt, float v, float vp, float ap) {
        const float k2f = xt.y + ksqyz;
        const float rs = rsqrtf(k2f);
        int kb = bin_fast(b, k2f * rs);
        if (kb == -2) kb = bin_exact(b, GS ? a.ax.ksq[0][i] : ksqx[i], ksqy, ksqz, kz);
        if (dcrow && i == 0) kb = -1;   // DC mode skipped
        kb -= a.kb0;
        if (kb < 0 || kb >= a.nbw) return;
        // the partner k_x = -i (none for i = 0 and the Nyquist index M)
        const bool pair = i != 0 && i != M;
        if (!pair) {
          vp = 0.f;
          ap = 0.f;
        }
        const float w = v * (xt.x * wyz);
        const float wp = vp * (ap * wyz);
        if (EPI == PSH_EPI_SIM) {
          const float kvx = (float)i * a.vec0f;
          const float mu = (kvx * b.losf[0] + dyz) * rs;
          const float mup = (-kvx * b.losf[0] + dyz) * rs;
          // accumulator of pole slot l, bin kb and |mu| cell c:
          // (l nbw + kb) nmu + c (nmu = 1: no cells)
          const int nmu = b.nmu;
          const int stride = a.nbw * nmu;
          int cm = 0, cp = 0;
          if (nmu > 1) {
            cm = min((int)(fabsf(mu) * (float)nmu), nmu - 1);
            cp = min((int)(fabsf(mup) * (float)nmu), nmu - 1);
          }
          const int base = kb * nmu + cm;
          if (a.even024) {
            // L_0 = 1, L_2 and L_4 by Horner in mu^2
            const float x2 = mu * mu;
            const float p2 = __builtin_fmaf(1.5f, x2, -0.5f);
            const float p4 = __builtin_fmaf(__builtin_fmaf(4.375f, x2, -3.75f), x2, 0.375f);
            if (a.losx0) {   // both modes of the pair share mu, and so the cell
              const float s0 = w + wp;
              acc_fixed(&ibins[base], s0, sdown);
              acc_fixed(&ibins[stride + base], s0 * p2, sdown);
              acc_fixed(&ibins[2 * stride + base], s0 * p4, sdown);
              return;
            }
            const float y2 = mup * mup;
            const float q2 = __builtin_fmaf(1.5f, y2, -0.5f);
            const float q4 = __builtin_fmaf(__builtin_fmaf(4.375f, y2, -3.75f), y2, 0.375f);
            if (cm == cp) {
              acc_fixed(&ibins[base], w + wp, sdown);
              acc_fixed(&ibins[stride + base], w * p2 + wp * q2, sdown);
              acc_fixed(&ibins[2 * stride + base], w * p4 + wp * q4, sdown);
            } else {
              const int basep = kb * nmu + cp;
              acc_fixed(&ibins[base], w, sdown);
              acc_fixed(&ibins[stride + base], w * p2, sdown);
              acc_fixed(&ibins[2 * stride + base], w * p4, sdown);
              acc_fixed(&ibins[basep], wp, sdown);
              acc_fixed(&ibins[stride + basep], wp * q2, sdown);
              acc_fixed(&ibins[2 * stride + basep], wp * q4, sdown);
            }
            return;
          }
          const float x2 = mu * mu, y2 = mup * mup;
          const int basep = kb * nmu + cp;
          // the general pole set (rare): a rolled loop keeps its registers
          // out of the common path's
#pragma unroll 1
          for (int l = 0; l < PSH_MAX_NL; l++) {
            if (l >= b.nl) break;
            if (!((lmask >> l) & 1)) continue;
            const float4 cf = legl[l];
            float pl = ((cf.w * x2 + cf.z) * x2 + cf.y) * x2 + cf.x;
            float ql = ((cf.w * y2 + cf.z) * y2 + cf.y) * y2 + cf.x;
            if ((a.oddmask >> l) & 1) {
              pl *= mu;
              ql *= mup;
            }
            if (cm == cp) {
              acc_fixed(&ibins[l * stride + base], w * pl + wp * ql, sdown);
            } else {
              acc_fixed(&ibins[l * stride + base], w * pl, sdown);
              acc_fixed(&ibins[l * stride + basep], wp * ql, sdown);
            }
          }
        } else {   // SURVEY0, and SURVEYL, whose Y_lm weights are in v already
          acc_fixed(&ibins[kb], w + wp, sdown);
        }
      };
      // The next pair's operands are read before the current pair's
      // accumulations: the compiler cannot tell the accumulators from the
      // tile and would otherwise wait for each pair's atomics before
      // reading the next pair.  Partner indices outside the pairs (i = 0,
      // M) read a valid entry that mode_pair discards.
      auto fetch = [&](int k, float4 &xt, float &v, float &vp, float &ap) {
        const int i = i0 + k * S;
        const int ip = (i != 0 && i != M) ? N - i : i;
        xt = xtab[i];
        v = vrow[L::off(k * S)];
        vp = lds[L::idx(c, ip)];
        ap = xtab[ip].x;
      };
      const int kend = i0 <= M ? (M - i0) / S : -1;   // last k with i <= M
      float4 xt = make_float4(0.f, 0.f, 0.f, 0.f);
      float v = 0.f, vp = 0.f, ap = 0.f;
      if (kend >= 0) fetch(0, xt, v, vp, ap);
#pragma unroll 1
      for (int k = 0; k <= kend; k++) {
        float4 xtn = xt;
        float vn = 0.f, vpn = 0.f, apn = 0.f;
        if (k < kend) fetch(k + 1, xtn, vn, vpn, apn);
        mode_pair(i0 + k * S, xt, v, vp, ap);
        xt = xtn;
        v = vn;
        vp = vpn;
        ap = apn;
      }
    }
    __syncthreads();
    // fold the tile's integer sums into the FP64 sums (each accumulator by
    // the thread that also writes it out), then clear them for the next tile
    const double sup = ldexp(1.0, E);
    for (int q2 = threadIdx.x; q2 < a.nacc; q2 += T) {
      const long long t = ibins[q2];
      if (t) {
        if (GS) out[out_idx(q2)] += (double)t * sup;
        else bins[q2] += (double)t * sup;
        ibins[q2] = 0;
      }
    }
    if (threadIdx.x == 0) *tmaxp = 0u;
    __syncthreads();
    tile = next;
  }
  // the window's accumulators into the workgroup's row of all bins
  if (!GS)
    for (int i = threadIdx.x; i < a.nacc; i += T) out[out_idx(i)] = bins[i];
}

// ---- dispatch ------------------------------------------------------------------

template <template <int> class F, class... Args>
hipError_t dispatch_n(int N, Args &&...args) {
  switch (N) {
    case 4: return F<4>::run(args...);
    case 8: return F<8>::run(args...);
    case 16: return F<16>::run(args...);
    case 32: return F<32>::run(args...);
    case 64: return F<64>::run(args...);
    case 128: return F<128>::run(args...);
    case 256: return F<256>::run(args...);
    case 512: return F<512>::run(args...);
    case 1024: return F<1024>::run(args...);
    case 2048: return F<2048>::run(args...);
    case 48: return F<48>::run(args...);
    case 96: return F<96>::run(args...);
    case 192: return F<192>::run(args...);
    case 384: return F<384>::run(args...);
    case 768: return F<768>::run(args...);
    case 1536: return F<1536>::run(args...);
    case 80: return F<80>::run(args...);
    case 160: return F<160>::run(args...);
    case 320: return F<320>::run(args...);
    case 640: return F<640>::run(args...);
    case 1280: return F<1280>::run(args...);
    default: return hipErrorInvalidValue;
  }
}

template <int N>
struct ZR2C {
  static hipError_t run(const float *in, float2 *out, int nmesh, size_t mstride,
                        const float2 *tw, const YlmDev &y, hipStream_t s, int kzlim) {
    const size_t nlines = (size_t)N * N * nmesh;
    const unsigned blocks = (unsigned)((nlines + ZCfg<N>::C - 1) / ZCfg<N>::C);
    const int H = N / 2 + PSH_PAD;
    auto go = [&](auto ntc) {
      constexpr bool NT = decltype(ntc)::value;
      if (y.ell > 0)
        hipLaunchKernelGGL((z_r2c_kernel<N, true, NT>), dim3(blocks), dim3(ZCfg<N>::T), 0,
                           s, in, out, nlines, mstride, H, tw, y, kzlim);
      else
        hipLaunchKernelGGL((z_r2c_kernel<N, false, NT>), dim3(blocks), dim3(ZCfg<N>::T), 0,
                           s, in, out, nlines, mstride, H, tw, y, kzlim);
      return hipGetLastError();
    };
    if constexpr (stream_mesh<N>()) return go(std::true_type{});
    else
      return stream_bytes(sizeof(float) * mstride * nmesh) ? go(std::true_type{})
                                                           : go(std::false_type{});
  }
};

template <int N>
struct ZC2R {
  static hipError_t run(float2 *data, int nmesh, size_t mstride, const float2 *tw,
                        float scale, hipStream_t s) {
    const size_t nlines = (size_t)N * N * nmesh;
    const unsigned blocks = (unsigned)((nlines + ZCfg<N>::C - 1) / ZCfg<N>::C);
    hipLaunchKernelGGL((z_c2r_kernel<N>), dim3(blocks), dim3(ZCfg<N>::T), 0, s, data,
                       nlines, mstride, N / 2 + PSH_PAD, tw, scale);
    return hipGetLastError();
  }
};

template <int N>
struct Col {
  template <int S>
  static hipError_t run_s(const float2 *in, float2 *out, int nmesh, size_t mstride, int axis,
                          const float2 *tw, int inverse, hipStream_t s, const int *ylive) {
    if constexpr (!ColCfg<N, S>::OK) {
      return hipErrorInvalidValue;
    } else {
      constexpr int C = ColCfg<N, S>::C;
      constexpr int CX = XCfg<N, true>::C;
      static_assert(CX == XCfg<N, false>::C && (CX & (CX - 1)) == 0, "one x tile width");
      int cxs = 0;
      while ((1 << cxs) < CX) cxs++;
      const int H = N / 2 + PSH_PAD;
      const int ntz = (N / 2 + 1 + C - 1) / C;
      const size_t OS = axis == 1 ? (size_t)N * H : (size_t)H;
      const size_t ES = axis == 1 ? (size_t)H : (size_t)N * H;
      const unsigned blocks = (unsigned)((size_t)N * ntz * nmesh);
      constexpr size_t shmem = ColCfg<N, S>::LDS;
      static_assert(shmem <= 160 * 1024, "column tile exceeds the LDS");
      auto go = [&](auto ntc) {
        constexpr bool NT = decltype(ntc)::value;
        auto k = inverse ? col_kernel<N, true, NT, S> : col_kernel<N, false, NT, S>;
        if (shmem > 64 * 1024)
          (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)shmem);
        hipLaunchKernelGGL(k, dim3(blocks), dim3(ColCfg<N, S>::T), shmem, s, in, out, ntz, OS,
                           ES, mstride / 2, tw, ylive, cxs);
        return hipGetLastError();
      };
      if constexpr (stream_mesh<N>()) return go(std::true_type{});
      else
        return stream_bytes(sizeof(float) * mstride * nmesh) ? go(std::true_type{})
                                                             : go(std::false_type{});
    }
  }
  static hipError_t run(const float2 *in, float2 *out, int nmesh, size_t mstride, int axis,
                        const float2 *tw, int inverse, hipStream_t s, const int *ylive,
                        int sched) {
    switch (sched) {
      case 1: return run_s<1>(in, out, nmesh, mstride, axis, tw, inverse, s, ylive);
      case 2: return run_s<2>(in, out, nmesh, mstride, axis, tw, inverse, s, ylive);
      case 3: return run_s<3>(in, out, nmesh, mstride, axis, tw, inverse, s, ylive);
      case 4: return run_s<4>(in, out, nmesh, mstride, axis, tw, inverse, s, ylive);
      case 5: return run_s<5>(in, out, nmesh, mstride, axis, tw, inverse, s, ylive);
      default: return run_s<0>(in, out, nmesh, mstride, axis, tw, inverse, s, ylive);
    }
  }
  static int ok(int sched) {
    switch (sched) {
      case 0: return ColCfg<N, 0>::OK;
      case 1: return ColCfg<N, 1>::OK;
      case 2: return ColCfg<N, 2>::OK;
      case 3: return ColCfg<N, 3>::OK;
      case 4: return ColCfg<N, 4>::OK;
      case 5: return ColCfg<N, 5>::OK;
      default: return 0;
    }
  }
};

template <int N>
struct XBin {
  using KFn = void (*)(XBinDev, int, int, const float2 *);
  // The kernel of a window of nacc accumulators, its dynamic LDS and
  // threads.  Two workgroups per CU (which the register budget allows up to
  // 512 threads, the plain pass up to N_g = 512) need at most half the LDS
  // each: when only the GS layout fits in that half, it is used (measured on
  // the learned correction's plain N_g = 512 pass with 8 |mu| cells: one
  // workgroup per CU otherwise).
  template <bool NT>
  static KFn pick(int epi, bool il, int nacc, size_t &shmem, int &T) {
    const size_t stat = il ? xbin_static_lds<N, true>() : xbin_static_lds<N, false>();
    const size_t np = (size_t)((nacc + 1) & ~1);
    T = il ? XCfg<N, true>::T : XCfg<N, false>::T;
    shmem = kXbinAccBytes * np + stat;
    bool gs = false;
    const size_t shmem_gs = kXbinAccBytesG * np + xbin_static_lds<N, false, true>();
    if (!il && T <= 512 && shmem > kXbinLdsMax / 2 && shmem_gs <= kXbinLdsMax / 2) {
      gs = true;
      shmem = shmem_gs;
    }
    KFn k = nullptr;
#define PSH_XB(E, I)                                                 \
  {                                                                     \
    k = x_bin_kernel<N, E, I, NT, false>;                               \
    if (!(I) && gs) k = x_bin_kernel<N, E, false, NT, true>;            \
  }
    if (epi == PSH_EPI_SIM) {
      if (il) PSH_XB(PSH_EPI_SIM, true) else PSH_XB(PSH_EPI_SIM, false)
    } else if (epi == PSH_EPI_SURVEY0) {
      if (il) PSH_XB(PSH_EPI_SURVEY0, true) else PSH_XB(PSH_EPI_SURVEY0, false)
    } else {
      PSH_XB(PSH_EPI_SURVEYL, false)
    }
#undef PSH_XB
    if (shmem > 64 * 1024)
      (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shmem);
    return k;
  }

  static hipError_t run(const XBinDev &d, int epi, bool il, int grid, int nmock,
                        const float2 *tw, hipStream_t s) {
    const int C = il ? XCfg<N, true>::C : XCfg<N, false>::C;
    const int H = N / 2 + PSH_PAD;
    const int ntz = (N / 2 + 1 + C - 1) / C;
    const XBinDev &dd = d;
    auto go = [&](auto ntc) {
      constexpr bool NT = decltype(ntc)::value;
      size_t shmem;
      int T;
      KFn k = pick<NT>(epi, il, d.nacc, shmem, T);
      if (shmem > kXbinLdsMax) return hipErrorInvalidValue;
      hipLaunchKernelGGL(k, dim3(grid, nmock), dim3(T), shmem, s, dd, ntz, H, tw);
      return hipGetLastError();
    };
    // the x pass reads nmock fields (two when interlaced) of mstride_c values
    const size_t bytes = sizeof(float2) * d.mstride_c * (size_t)nmock * (il ? 2 : 1);
    if constexpr (stream_mesh<N>()) return go(std::true_type{});
    else return stream_bytes(bytes) ? go(std::true_type{}) : go(std::false_type{});
  }
};

// Workgroups per CU of the x pass kernel for a window of nacc accumulators.
template <int N>
struct XBinOcc {
  static hipError_t run(int epi, bool il, int nacc, int *per_cu) {
    size_t shmem;
    int T;
    auto k = XBin<N>::template pick<true>(epi, il, nacc, shmem, T);
    if (shmem > kXbinLdsMax) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, (const void *)k, T, shmem);
  }
};

}  // namespace

hipError_t psh_fft_twiddles_fill(int N, float2 *d_tw) {
  std::vector<float2> h(N);
  for (int m = 0; m < N; m++) {
    double ang = -2.0 * M_PI * (double)m / (double)N;
    h[m] = make_float2((float)cos(ang), (float)sin(ang));
  }
  return hipMemcpy(d_tw, h.data(), sizeof(float2) * N, hipMemcpyHostToDevice);
}

hipError_t psh_fft_twiddles(int N, float2 **d_tw) {
  hipError_t e = hipMalloc((void **)d_tw, sizeof(float2) * N);
  if (e != hipSuccess) return e;
  return psh_fft_twiddles_fill(N, *d_tw);
}

hipError_t psh_fft_z_r2c(int Ng, const float *in, float2 *out, int nmesh,
                         size_t mstride_floats, const float2 *tw,
                         const PshYlmR &ylm, hipStream_t s, int kzlim) {
  YlmDev y{};
  y.ell = ylm.ell;
  for (int a = 0; a < 3; a++) {
    y.org[a] = ylm.org[a];
    y.step[a] = ylm.step[a];
  }
  if (ylm.ell > 0) {
    if (ylm.ell > 6 || ylm.m < -ylm.ell || ylm.m > ylm.ell) return hipErrorInvalidValue;
    y.poly[0] = make_ylm_poly(ylm.ell, ylm.m, ylm.norm);
  }
  return dispatch_n<ZR2C>(Ng, in, out, nmesh, mstride_floats, tw, y, s, kzlim);
}

hipError_t psh_fft_z_c2r(int Ng, float2 *inout, int nmesh, size_t mstride_floats,
                         const float2 *tw, float scale, hipStream_t s) {
  return dispatch_n<ZC2R>(Ng, inout, nmesh, mstride_floats, tw, scale, s);
}

hipError_t psh_fft_y(int Ng, float2 *data, int nmesh, size_t mstride_floats,
                     const float2 *tw, int inverse, hipStream_t s, const int *ylive,
                     int sched) {
  return dispatch_n<Col>(Ng, (const float2 *)data, data, nmesh, mstride_floats, 1, tw,
                         inverse, s, inverse ? nullptr : ylive, sched);
}

hipError_t psh_fft_x(int Ng, const float2 *in, float2 *out, int nmesh,
                     size_t mstride_floats, const float2 *tw, int inverse,
                     hipStream_t s) {
  return dispatch_n<Col>(Ng, in, out, nmesh, mstride_floats, 0, tw, inverse, s,
                         (const int *)nullptr, 0);
}

namespace {
template <int N>
struct ColOk {
  static hipError_t run(int sched, int *ok) {
    *ok = Col<N>::ok(sched);
    return hipSuccess;
  }
};
}  // namespace

static_assert(kColScheds == PSH_FFT_Y_SCHEDS, "schedule count");

int psh_fft_y_sched_ok(int Ng, int sched) {
  int ok = 0;
  if (sched < 0 || sched >= kColScheds) return 0;
  return dispatch_n<ColOk>(Ng, sched, &ok) == hipSuccess && ok;
}

int psh_xbin_nacc(const PshBins &b, int epi) {
  return epi == PSH_EPI_SIM ? b.nl * b.nbin * b.nmu : b.nbin;
}

int psh_xbin_tile_cols(int Ng) { return Ng > 1024 ? XCfg<2048, true>::C : XCfg<1024, true>::C; }

namespace {
template <int N>
struct XBinLds {
  static size_t run(bool il, int nacc) {
    return kXbinAccBytes * ((nacc + 1) & ~1) +
           (il ? xbin_static_lds<N, true>() : xbin_static_lds<N, false>());
  }
};
}  // namespace

extern "C" int psh_fft_plan(int Ng, int plan, int *radices, int max) {
  if (Ng < 2 || Ng > 4096 || plan < 0 || plan > 2 || !radices) return 0;
  const int odd = odd_part(Ng);
  if ((odd != 1 && odd != 3 && odd != 5) || (odd > 1 && pow2_part(Ng) < 2)) return 0;
  const int n = npasses(Ng, plan);
  for (int p = 0; p < n && p < max; p++) radices[p] = radix(Ng, p, plan);
  return n;
}

bool psh_grid_supported(int Ng) {
  switch (Ng) {
    case 4: case 8: case 16: case 32: case 64: case 128: case 256: case 512: case 1024:
    case 2048:
    case 48: case 96: case 192: case 384: case 768: case 1536:
    case 80: case 160: case 320: case 640: case 1280:
      return true;
    default:
      return false;
  }
}

size_t psh_xbin_lds_bytes(int Ng, bool il, int nacc) {
  switch (Ng) {
    case 4: return XBinLds<4>::run(il, nacc);
    case 8: return XBinLds<8>::run(il, nacc);
    case 16: return XBinLds<16>::run(il, nacc);
    case 32: return XBinLds<32>::run(il, nacc);
    case 64: return XBinLds<64>::run(il, nacc);
    case 128: return XBinLds<128>::run(il, nacc);
    case 256: return XBinLds<256>::run(il, nacc);
    case 512: return XBinLds<512>::run(il, nacc);
    case 1024: return XBinLds<1024>::run(il, nacc);
    case 2048: return XBinLds<2048>::run(il, nacc);
    case 48: return XBinLds<48>::run(il, nacc);
    case 96: return XBinLds<96>::run(il, nacc);
    case 192: return XBinLds<192>::run(il, nacc);
    case 384: return XBinLds<384>::run(il, nacc);
    case 768: return XBinLds<768>::run(il, nacc);
    case 1536: return XBinLds<1536>::run(il, nacc);
    case 80: return XBinLds<80>::run(il, nacc);
    case 160: return XBinLds<160>::run(il, nacc);
    case 320: return XBinLds<320>::run(il, nacc);
    case 640: return XBinLds<640>::run(il, nacc);
    case 1280: return XBinLds<1280>::run(il, nacc);
    default: return 0;
  }
}

int psh_xbin_window_bins(int Ng, bool il, int per_bin) {
  const size_t fixed = psh_xbin_lds_bytes(Ng, il, 0);
  if (fixed == 0 || fixed >= kXbinLdsMax || per_bin < 1) return 0;
  long nb = (long)((kXbinLdsMax - fixed) / kXbinAccBytes) / per_bin;
  while (nb > 0 && psh_xbin_lds_bytes(Ng, il, (int)(nb * per_bin)) > kXbinLdsMax) nb--;
  return nb > (1 << 30) ? (1 << 30) : (int)nb;
}

int psh_xbin_per_cu(int Ng, int epi, bool il, int nacc) {
  int nb = 0;
  if (dispatch_n<XBinOcc>(Ng, epi, il, nacc, &nb) != hipSuccess) return 0;
  return nb;
}

hipError_t psh_fft_x_bin(int Ng, const PshXBinArgs &a, const float2 *tw,
                         hipStream_t s) {
  XBinDev d{};
  d.A = a.A;
  d.B = a.B;
  d.F0 = a.F0;
  d.store = a.store;
  d.mstride_c = a.mstride_floats / 2;
  d.ax = a.ax;
  d.bins = a.bins;
  d.phase = a.phase;
  d.partials = a.partials;
  d.nacc_full = psh_xbin_nacc(a.bins, a.epi);
  const int per_bin = a.epi == PSH_EPI_SIM ? a.bins.nl * a.bins.nmu : 1;
  d.kb0 = a.nbw > 0 ? a.kb0 : 0;
  d.nbw = a.nbw > 0 ? a.nbw : a.bins.nbin;
  if (d.kb0 < 0 || d.kb0 + d.nbw > a.bins.nbin) return hipErrorInvalidValue;
  d.nacc = per_bin * d.nbw;
  // bound of a pair's |contribution| over the tile's largest |v| w_yz
  // (which phase 1 measures; for l > 0 survey fields v carries the Y_lm
  // weights), times N C (at least 1024): the largest x alias factor and
  // |L_l| below 1.01, two modes per pair, a 2 % margin for FP32 rounding
  {
    const int C = psh_xbin_tile_cols(Ng);
    double w = (double)a.ax.almax[0] * 1.01;
    w *= 2.0 * 1.02;
    const double nc = (double)Ng * C;
    d.xbound = w * (nc > 1024.0 ? nc : 1024.0);
  }
  d.losx0 = a.bins.losf[0] == 0.f;
  d.has_dc = a.has_dc;
  d.dcf = (float)a.dc;
  d.kd0f = (float)((double)a.vec0f / (2 * M_PI));
  d.nf = 1;
  if (a.epi == PSH_EPI_SURVEYL) {
    if (a.nf < 1 || a.nf > PSH_MAX_YLM_FIELDS) return hipErrorInvalidValue;
    d.nf = a.nf;
    d.fstride_c = a.fstride_floats / 2;
    for (int f = 0; f < a.nf; f++) {
      const int ell = a.ells[f], m = a.ms[f];
      if (ell < 1 || ell > 6 || m < -ell || m > ell) return hipErrorInvalidValue;
      d.yp[f] = make_ylm_poly(ell, m, a.ynorms[f]);
    }
  }
  d.vec0f = a.vec0f;
  d.ylive = a.ylive;
  // Legendre polynomials L_0..L_6 (math/legpoly.h:47-61 of the reference) as
  // polynomials in mu^2
  static const float kLeg[7][4] = {
      {1.f, 0.f, 0.f, 0.f},         {1.f, 0.f, 0.f, 0.f},
      {-0.5f, 1.5f, 0.f, 0.f},      {-1.5f, 2.5f, 0.f, 0.f},
      {0.375f, -3.75f, 4.375f, 0.f}, {1.875f, -8.75f, 7.875f, 0.f},
      {-0.3125f, 6.5625f, -19.6875f, 14.4375f}};
  d.oddmask = 0;
  for (int l = 0; l < a.bins.nl && l < PSH_MAX_NL; l++) {
    const int ell = a.bins.poles[l];
    if (ell < 0 || ell > 6) return hipErrorInvalidValue;
    for (int t = 0; t < 4; t++) d.legc[l][t] = kLeg[ell][t];
    if (ell & 1) d.oddmask |= 1 << l;
  }
  d.even024 = a.bins.nl == 3 && a.bins.poles[0] == 0 && a.bins.poles[1] == 2 &&
              a.bins.poles[2] == 4;
  return dispatch_n<XBin>(Ng, d, a.epi, a.B != nullptr, a.grid, a.nmock, tw, s);
}
