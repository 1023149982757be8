// k_mlp.hip -- the learned alias correction: a 4-layer perceptron (inputs
// padded to 32, three hidden layers of 64 with ReLU, one output) trained and
// evaluated on gfx950 matrix cores: every product -- forward, weight
// gradients H^T dZ and input gradients dZ W^T -- runs as bf16 MFMA
// (v_mfma_f32_16x16x16_bf16) with FP32 accumulation; FP32 master weights,
// Adam updates.  No counterpart in the reference.
#include <math.h>

#include <hip/hip_runtime.h>

#include "powspec_hip.h"
#include "psh_internal.h"

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int IN = 32, H = 64, OUT = 16;   // padded widths
// parameter layout (FP32): W1[IN][H] b1[H] W2[H][H] b2[H] W3[H][H] b3[H] W4[H][OUT] b4[OUT]
constexpr int oW1 = 0, ob1 = oW1 + IN * H, oW2 = ob1 + H, ob2 = oW2 + H * H, oW3 = ob2 + H,
              ob3 = oW3 + H * H, oW4 = ob3 + H, ob4 = oW4 + H * OUT, NPAR = ob4 + OUT;

__device__ inline short bf16(float x) {   // round to nearest even
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// C(r, n) = sum_k A(r, k) B(k, n) for the wave's 16 rows, K and N multiples
// of 16.  Operand element (i, j) at base[i * s0 + j * s1] (FP32, converted
// to bf16 on load), so transposed operands cost nothing.  Lane layout of
// the 16x16x16 MFMA: A[l%16][4(l/16)+j], B[4(l/16)+j][l%16],
// C[4(l/16)+i][l%16].
template <int K, int N>
__device__ inline void wave_mma(const float *A, int a0, int a1, const float *B, int b0, int b1,
                                float *C, int c0, int c1) {
  const int l = (int)(threadIdx.x & 63);
  const int r = l & 15, q = l >> 4;
#pragma unroll
  for (int nt = 0; nt < N / 16; nt++) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < K / 16; kt++) {
      s4 a, b;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int k = kt * 16 + 4 * q + j;
        a[j] = bf16(A[r * a0 + k * a1]);
        b[j] = bf16(B[k * b0 + (nt * 16 + r) * b1]);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) C[(4 * q + i) * c0 + (nt * 16 + r) * c1] = acc[i];
  }
}

// bias + optional ReLU on a wave's [16][n] LDS block
__device__ inline void bias_act(float *Z, const float *b, int n, bool relu) {
  const int l = (int)(threadIdx.x & 63);
  for (int e = l; e < 16 * n; e += 64) {
    float v = Z[e] + b[e % n];
    Z[e] = relu ? fmaxf(v, 0.f) : v;
  }
}

struct WaveBufs {
  float x[16 * IN], h1[16 * H], h2[16 * H], h3[16 * H], y[16 * OUT];
  float d1[16 * H], d2[16 * H], t[64 * OUT];
};

// forward of the wave's 16 samples; rows past B are zero
__device__ inline void forward(const float *P, const float *X, int s0, int B, WaveBufs &w) {
  const int l = (int)(threadIdx.x & 63);
  for (int e = l; e < 16 * IN; e += 64) {
    const int s = s0 + e / IN;
    w.x[e] = s < B ? X[(size_t)s * IN + e % IN] : 0.f;
  }
  __builtin_amdgcn_wave_barrier();
  wave_mma<IN, H>(w.x, IN, 1, P + oW1, H, 1, w.h1, H, 1);
  __builtin_amdgcn_wave_barrier();
  bias_act(w.h1, P + ob1, H, true);
  __builtin_amdgcn_wave_barrier();
  wave_mma<H, H>(w.h1, H, 1, P + oW2, H, 1, w.h2, H, 1);
  __builtin_amdgcn_wave_barrier();
  bias_act(w.h2, P + ob2, H, true);
  __builtin_amdgcn_wave_barrier();
  wave_mma<H, H>(w.h2, H, 1, P + oW3, H, 1, w.h3, H, 1);
  __builtin_amdgcn_wave_barrier();
  bias_act(w.h3, P + ob3, H, true);
  __builtin_amdgcn_wave_barrier();
  wave_mma<H, OUT>(w.h3, H, 1, P + oW4, OUT, 1, w.y, OUT, 1);
  __builtin_amdgcn_wave_barrier();
  bias_act(w.y, P + ob4, OUT, false);
  __builtin_amdgcn_wave_barrier();
}

constexpr int kWaves = 4;

// ---- inference: weights resident in LDS as bf16 MFMA fragments -------------
//
// A workgroup converts the FP32 weights once into bf16 B fragments (for each
// 16x16 tile (kt, nt): lane l holds W[kt 16 + 4(l/16) + j][nt 16 + l%16],
// j = 0..3, as 8 contiguous bytes) and its waves then loop over tiles of 16
// samples.  Activations stay bf16 in a padded [16][H + 8] row block per wave;
// bias and ReLU are applied to the accumulators in registers.  The bf16
// roundings are the same as in forward(): every operand rounded to nearest
// even once, FP32 accumulation.
constexpr int kInferWaves = 8;
constexpr int ROWP = H + 8;   // padded bf16 row (bank spread of the A reads)
constexpr int FRAG1 = IN * H, FRAG2 = H * H, FRAG4 = H * OUT;

struct InferLds {
  short w1[FRAG1], w2[FRAG2], w3[FRAG2], w4[FRAG4];
  float b[3 * H + OUT];
  short act[kInferWaves][2][16 * ROWP];
};

template <int K, int N>
__device__ inline void pack_frags(const float *W, short *F) {
  for (int e = threadIdx.x; e < K * N; e += blockDim.x) {
    const int tile = e >> 8, l = (e >> 2) & 63, j = e & 3;
    const int kt = tile / (N / 16), nt = tile % (N / 16);
    F[e] = bf16(W[(kt * 16 + 4 * (l >> 4) + j) * N + nt * 16 + (l & 15)]);
  }
}

// acc tiles of C = A W for the wave's 16 rows: A bf16 rows [16][ROWP]
// (or, for the input layer, FP32 rows of X), W fragments F.
template <int K, int N, bool RELU>
__device__ inline void infer_layer(const short *A, const short *F, const float *bias,
                                   short *out) {
  const int l = (int)(threadIdx.x & 63), r = l & 15, q = l >> 4;
#pragma unroll
  for (int nt = 0; nt < N / 16; nt++) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < K / 16; kt++) {
      const s4 a = *reinterpret_cast<const s4 *>(A + r * ROWP + kt * 16 + 4 * q);
      const s4 b = *reinterpret_cast<const s4 *>(F + (kt * (N / 16) + nt) * 256 + 4 * l);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const float bn = bias[nt * 16 + r];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      float v = acc[i] + bn;
      if (RELU) v = fmaxf(v, 0.f);
      out[(4 * q + i) * ROWP + nt * 16 + r] = bf16(v);
    }
  }
}

// blockIdx.y: network of an ensemble (parameters P + y pstride, outputs
// Y + y B); every network reads the same rows X.
__global__ void __launch_bounds__(64 * kInferWaves)
mlp_infer_kernel(const float *__restrict__ P0, size_t pstride, const float *__restrict__ X,
                 int B, float *__restrict__ Y0) {
  const float *__restrict__ P = P0 + (size_t)blockIdx.y * pstride;
  float *__restrict__ Y = Y0 + (size_t)blockIdx.y * B;
  extern __shared__ __align__(16) unsigned char infer_raw[];
  InferLds &S = *reinterpret_cast<InferLds *>(infer_raw);
  pack_frags<IN, H>(P + oW1, S.w1);
  pack_frags<H, H>(P + oW2, S.w2);
  pack_frags<H, H>(P + oW3, S.w3);
  pack_frags<H, OUT>(P + oW4, S.w4);
  for (int e = threadIdx.x; e < 3 * H + OUT; e += blockDim.x)
    S.b[e] = e < H ? P[ob1 + e] : e < 2 * H ? P[ob2 + e - H] : e < 3 * H ? P[ob3 + e - 2 * H]
                                                                          : P[ob4 + e - 3 * H];
  __syncthreads();
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  short *a0 = S.act[wv][0], *a1 = S.act[wv][1];
  const int ntile = (B + 15) / 16;
  for (int t = blockIdx.x * kInferWaves + wv; t < ntile; t += gridDim.x * kInferWaves) {
    const int s0 = t * 16;
    // input rows -> bf16 [16][ROWP] (columns past IN unused)
#pragma unroll
    for (int kt = 0; kt < IN / 16; kt++) {
      const int s = s0 + r;
      float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
      if (s < B) x = *reinterpret_cast<const float4 *>(X + (size_t)s * IN + kt * 16 + 4 * q);
      short *d = a0 + r * ROWP + kt * 16 + 4 * q;
      d[0] = bf16(x.x);
      d[1] = bf16(x.y);
      d[2] = bf16(x.z);
      d[3] = bf16(x.w);
    }
    __builtin_amdgcn_wave_barrier();
    infer_layer<IN, H, true>(a0, S.w1, S.b, a1);
    __builtin_amdgcn_wave_barrier();
    infer_layer<H, H, true>(a1, S.w2, S.b + H, a0);
    __builtin_amdgcn_wave_barrier();
    infer_layer<H, H, true>(a0, S.w3, S.b + 2 * H, a1);
    __builtin_amdgcn_wave_barrier();
    // output layer: only column 0 is read (lanes with r == 0 hold rows 4q..4q+3)
    {
      f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kt = 0; kt < H / 16; kt++) {
        const s4 a = *reinterpret_cast<const s4 *>(a1 + r * ROWP + kt * 16 + 4 * q);
        const s4 b = *reinterpret_cast<const s4 *>(S.w4 + kt * 256 + 4 * l);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
      }
      if (r == 0) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int s = s0 + 4 * q + i;
          if (s < B) Y[s] = acc[i] + S.b[3 * H];
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Gradients are summed over the waves of a batch in 64-bit fixed point
// (value * 2^40): integer addition is associative, so a training run is
// reproducible bit for bit whatever order the waves finish in (FP32 atomics
// would make every run a different trajectory).  Parts below 2^-41 are
// dropped, far below what Adam's eps = 1e-8 resolves.  A wave's
// contribution is clamped to |g| <= 2^16 (NaN to the bound) so that the
// conversion is always defined; the 64-bit sum holds 2^7 saturated
// contributions, and those of a converging run are orders of magnitude
// smaller (the loss gradient carries the 1/B of the mean).
constexpr float kGradScale = 1099511627776.0f;           // 2^40
constexpr double kGradUnscale = 1.0 / 1099511627776.0;   // 2^-40
constexpr float kGradClamp = 65536.0f;                   // 2^16

__device__ __forceinline__ void grad_add(long long *g, float v) {
  const float c = fminf(fmaxf(v, -kGradClamp), kGradClamp);   // NaN -> clamp bound
  atomicAdd(reinterpret_cast<unsigned long long *>(g),
            (unsigned long long)__float2ll_rn(c * kGradScale));
}

// dW[m][n] += sum_s A[s][m] D[s][n] over the wave's 16 samples (M = rows of
// the weight), via MFMA with K = 16 samples; db[n] += sum_s D[s][n]
template <int M, int N>
__device__ inline void weight_grad(const float *A, int lda, const float *D, int ldd, long long *G,
                                   long long *gb, float *tmp) {
  const int l = (int)(threadIdx.x & 63);
#pragma unroll
  for (int mt = 0; mt < M / 16; mt++) {
    // tmp(r, n) = sum_s A[s][mt*16 + r] D[s][n]
    wave_mma<16, N>(A + mt * 16, 1, lda, D, ldd, 1, tmp, N, 1);
    __builtin_amdgcn_wave_barrier();
    for (int e = l; e < 16 * N; e += 64) grad_add(&G[(mt * 16 + e / N) * N + e % N], tmp[e]);
    __builtin_amdgcn_wave_barrier();
  }
  for (int n = l; n < N; n += 64) {
    float s = 0.f;
    for (int r = 0; r < 16; r++) s += D[r * ldd + n];
    grad_add(&gb[n], s);
  }
}

// loss = sum (y - t)^2 / B_total over the batch, gradient accumulated in G
__global__ void __launch_bounds__(64 * kWaves) mlp_grad_kernel(const float *P, const float *X,
                                                               const float *T, int B,
                                                               float inv_total, long long *G,
                                                               float *loss) {
  __shared__ WaveBufs wb[kWaves];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  WaveBufs &w = wb[wv];
  const int s0 = (blockIdx.x * kWaves + wv) * 16;
  forward(P, X, s0, B, w);
  // dY (only output 0 is trained) into y
  float lsum = 0.f;
  for (int e = l; e < 16 * OUT; e += 64) {
    const int s = s0 + e / OUT;
    float g = 0.f;
    if (e % OUT == 0 && s < B) {
      const float d = w.y[e] - T[s];
      lsum += d * d * inv_total;
      g = 2.f * d * inv_total;
    }
    w.y[e] = g;
  }
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
  if (l == 0) atomicAdd(loss, lsum);
  __builtin_amdgcn_wave_barrier();
  // layer 4
  weight_grad<H, OUT>(w.h3, H, w.y, OUT, G + oW4, G + ob4, w.t);
  wave_mma<OUT, H>(w.y, OUT, 1, P + oW4, 1, OUT, w.d1, H, 1);   // dH3 = dY W4^T
  __builtin_amdgcn_wave_barrier();
  for (int e = l; e < 16 * H; e += 64) w.d1[e] = w.h3[e] > 0.f ? w.d1[e] : 0.f;
  __builtin_amdgcn_wave_barrier();
  // layer 3
  weight_grad<H, H>(w.h2, H, w.d1, H, G + oW3, G + ob3, w.t);
  wave_mma<H, H>(w.d1, H, 1, P + oW3, 1, H, w.d2, H, 1);        // dH2 = dZ3 W3^T
  __builtin_amdgcn_wave_barrier();
  for (int e = l; e < 16 * H; e += 64) w.d2[e] = w.h2[e] > 0.f ? w.d2[e] : 0.f;
  __builtin_amdgcn_wave_barrier();
  // layer 2
  weight_grad<H, H>(w.h1, H, w.d2, H, G + oW2, G + ob2, w.t);
  wave_mma<H, H>(w.d2, H, 1, P + oW2, 1, H, w.d1, H, 1);        // dH1 = dZ2 W2^T
  __builtin_amdgcn_wave_barrier();
  for (int e = l; e < 16 * H; e += 64) w.d1[e] = w.h1[e] > 0.f ? w.d1[e] : 0.f;
  __builtin_amdgcn_wave_barrier();
  // layer 1
  weight_grad<IN, H>(w.x, IN, w.d1, H, G + oW1, G + ob1, w.t);
}

__global__ void adam_kernel(float *P, long long *G, float *m, float *v, int n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float g = (float)((double)G[i] * kGradUnscale);
    m[i] = b1 * m[i] + (1.f - b1) * g;
    v[i] = b2 * v[i] + (1.f - b2) * g * g;
    P[i] -= lr * (m[i] * c1) / (sqrtf(v[i] * c2) + eps);
    G[i] = 0;
  }
}

template <typename V>
__global__ void zero_kernel(V *p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = V(0);
}

__global__ void grad_out_kernel(const long long *G, float *out, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = (float)((double)G[i] * kGradUnscale);
}

unsigned blocks_for(int B) { return (unsigned)((B + 16 * kWaves - 1) / (16 * kWaves)); }

}  // namespace

extern "C" int psh_mlp_param_count(void) { return NPAR; }

namespace {

// Features of the learned alias correction, one row per (k bin, |mu| cell)
// of a plain (non-interlaced) box run, from the run's own raw cell sums
// S[l][b][c] (pole slot 0 the monopole):
//   [0]      k_b / k_Ny (bin centre)
//   [1]      centre of the |mu| cell, (c + 1/2) / nmu
//   [2..5]   one-hot assignment scheme
//   [6..21]  log10 of the un-subtracted monopole over the shot noise,
//            averaged over 16 equal bands of [k_Ny / 4, k_Ny] (an empty band
//            takes the populated bin nearest its centre): the spectrum's
//            shape towards the Nyquist frequency and its level
//   [22]     log10 of the bin's own un-subtracted monopole over the shot noise
//   [23]     log10(nmu S_0[b][c] / sum_c S_0[b][c]), clipped to [-2, 2]: the
//            cell's share of the bin (|mu| is uniform over directions)
// Bins without modes or power get zero rows (and no correction).
__device__ __forceinline__ double alias_kcen(int b, double kedge0, double dk, int logscale) {
  const double c = kedge0 + dk * (b + 0.5);
  return logscale ? pow(10.0, c) : c;
}

// Per bin: P[b] = the un-subtracted monopole over the shot noise (0 without
// modes or power) and S[b] = the bin's monopole cell sum.
__global__ void __launch_bounds__(256)
alias_bins_kernel(const double *__restrict__ cells, const unsigned long long *__restrict__ cnt,
                  int nbin, int nmu, double norm, double shot, double *__restrict__ P,
                  double *__restrict__ S) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= nbin) return;
  double s = 0.0;
  for (int c = 0; c < nmu; c++) s += cells[(size_t)b * nmu + c];
  S[b] = s;
  double p = 0.0;
  if (cnt[b]) {
    p = s / (norm * (double)cnt[b]) / shot;
    if (!(p > 0)) p = 0.0;
  }
  P[b] = p;
}

// The 16-band summary: 16 threads per band stride over the bins, each
// keeping its band sum, count and the populated bin nearest the band centre
// (the first in bin order on a tie); the threads' partials are combined in
// a fixed order.
__global__ void __launch_bounds__(256)
alias_bands_kernel(const double *__restrict__ P, int nbin, double kny, double kedge0, double dk,
                   int logscale, float *__restrict__ summ) {
  __shared__ double ssum[256], sbd[256];
  __shared__ int sn[256], sbest[256];
  const int t = threadIdx.x, band = t >> 4, lane = t & 15;
  const double e0 = kny * (0.25 + 0.75 * band / 16.0), e1 = kny * (0.25 + 0.75 * (band + 1) / 16.0);
  const double mid = 0.5 * (e0 + e1);
  double sum = 0.0, bestd = 1e300;
  int n = 0, best = -1;
  for (int b = lane; b < nbin; b += 16) {
    const double p = P[b];
    if (p <= 0) continue;
    const double k = alias_kcen(b, kedge0, dk, logscale);
    if (k >= e0 && k < e1) {
      sum += p;
      n++;
    }
    const double d = fabs(k - mid);
    if (d < bestd) {
      bestd = d;
      best = b;
    }
  }
  ssum[t] = sum;
  sn[t] = n;
  sbd[t] = bestd;
  sbest[t] = best;
  __syncthreads();
  if (lane == 0) {
    for (int q = 1; q < 16; q++) {
      sum += ssum[t + q];
      n += sn[t + q];
      const int bq = sbest[t + q];
      if (bq >= 0 && (sbd[t + q] < bestd || (sbd[t + q] == bestd && bq < best))) {
        bestd = sbd[t + q];
        best = bq;
      }
    }
    const double v = n ? sum / n : (best >= 0 ? P[best] : 1.0);
    summ[band] = (float)log10(v);
  }
}

// One thread per (bin, cell) row.
__global__ void __launch_bounds__(256)
alias_features_kernel(const double *__restrict__ cells, const double *__restrict__ P,
                      const double *__restrict__ S, const float *__restrict__ summ, int nbin,
                      int nmu, double kny, double kedge0, double dk, int logscale, int scheme,
                      float *__restrict__ X) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= nbin * nmu) return;
  const int b = r / nmu, c = r - b * nmu;
  float *x = X + (size_t)r * IN;
  for (int q = 0; q < IN; q++) x[q] = 0.f;
  const double p = P[b];
  if (p <= 0) return;
  const double sb = S[b];
  const double sc = cells[(size_t)b * nmu + c];
  x[0] = (float)(alias_kcen(b, kedge0, dk, logscale) / kny);
  x[1] = (float)((c + 0.5) / nmu);
  x[2 + (scheme & 3)] = 1.f;
  for (int q = 0; q < 16; q++) x[6 + q] = summ[q];
  x[22] = (float)log10(p);
  const double share = (sb > 0 && sc > 0) ? log10(nmu * sc / sb) : -2.0;
  x[23] = (float)fmin(2.0, fmax(-2.0, share));
}

// Corrected raw sums: out[l][b] = sum_c exp(f[b][c]) S[l][b][c], f the mean
// over the nnet networks' outputs (rows without features keep factor 1).
__global__ void __launch_bounds__(256)
alias_apply_kernel(const float *__restrict__ Y, int nnet, int rows, const float *__restrict__ X,
                   const double *__restrict__ cells, int nl, int nbin, int nmu,
                   double *__restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nl * nbin) return;
  const int l = q / nbin, b = q - l * nbin;
  double s = 0.0;
  for (int c = 0; c < nmu; c++) {
    const int r = b * nmu + c;
    double f = 0.0;
    if (X[(size_t)r * IN] != 0.f) {
      for (int n = 0; n < nnet; n++) f += Y[(size_t)n * rows + r];
      f /= nnet;
    }
    s += exp(f) * cells[((size_t)l * nbin + b) * nmu + c];
  }
  out[q] = s;
}

}  // namespace

size_t psh_alias_scratch_doubles(int nbin) { return 2 * (size_t)nbin + 8; }

hipError_t psh_alias_features(const double *cells, const unsigned long long *cnt, int nbin,
                              int nmu, double norm, double shot, double kny, double kedge0,
                              double dk, int logscale, int scheme, float *X, double *scratch,
                              hipStream_t s) {
  if (nbin < 1 || nmu < 1) return hipErrorInvalidValue;
  double *P = scratch, *S = scratch + nbin;
  float *summ = reinterpret_cast<float *>(scratch + 2 * (size_t)nbin);
  hipLaunchKernelGGL(alias_bins_kernel, dim3((nbin + 255) / 256), dim3(256), 0, s, cells, cnt,
                     nbin, nmu, norm, shot, P, S);
  hipLaunchKernelGGL(alias_bands_kernel, dim3(1), dim3(256), 0, s, P, nbin, kny, kedge0, dk,
                     logscale, summ);
  const int rows = nbin * nmu;
  hipLaunchKernelGGL(alias_features_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, cells, P,
                     S, summ, nbin, nmu, kny, kedge0, dk, logscale, scheme, X);
  return hipGetLastError();
}

hipError_t psh_alias_apply(const float *Y, int nnet, const float *X, const double *cells, int nl,
                           int nbin, int nmu, double *out, hipStream_t s) {
  const int n = nl * nbin;
  hipLaunchKernelGGL(alias_apply_kernel, dim3((n + 255) / 256), dim3(256), 0, s, Y, nnet,
                     nbin * nmu, X, cells, nl, nbin, nmu, out);
  return hipGetLastError();
}

hipError_t psh_mlp_forward_async(const float *d_params, const float *d_X, int B, float *d_Y,
                                 hipStream_t s, int nnet) {
  if (B <= 0 || nnet <= 0) return hipSuccess;
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const int ntile = (B + 15) / 16;
  int grid = (ntile + kInferWaves - 1) / kInferWaves;
  if (grid > 2 * ncu) grid = 2 * ncu;
  const size_t lds = sizeof(InferLds);
  (void)hipFuncSetAttribute((const void *)mlp_infer_kernel,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(mlp_infer_kernel, dim3(grid, nnet), dim3(64 * kInferWaves), lds, s,
                     d_params, (size_t)psh_mlp_param_count(), d_X, B, d_Y);
  return hipGetLastError();
}

extern "C" int psh_mlp_forward(const float *d_params, const float *d_X, int B, float *d_Y) {
  if (!d_params || !d_X || !d_Y || B < 0) return PSH_ERR_ARG;
  if (B == 0) return PSH_OK;
  // persistent: at most two workgroups per CU, each packing the weights once
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const int ntile = (B + 15) / 16;
  int grid = (ntile + kInferWaves - 1) / kInferWaves;
  if (grid > 2 * ncu) grid = 2 * ncu;
  const size_t lds = sizeof(InferLds);
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void *)mlp_infer_kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    attr = true;
  }
  hipLaunchKernelGGL(mlp_infer_kernel, dim3(grid), dim3(64 * kInferWaves), lds, nullptr,
                     d_params, (size_t)0, d_X, B, d_Y);
  return hipDeviceSynchronize() == hipSuccess ? PSH_OK : PSH_ERR_DEVICE;
}

extern "C" int psh_mlp_grad(const float *d_params, const float *d_X, const float *d_T, int B,
                            float *d_grads, float *loss) {
  if (!d_params || !d_X || !d_T || !d_grads || B < 1) return PSH_ERR_ARG;
  float *d_loss = nullptr;
  long long *G = nullptr;
  if (hipMalloc(&d_loss, sizeof(float)) != hipSuccess) return PSH_ERR_MEMORY;
  if (hipMalloc(&G, sizeof(long long) * NPAR) != hipSuccess) {
    (void)hipFree(d_loss);
    return PSH_ERR_MEMORY;
  }
  hipLaunchKernelGGL(zero_kernel<long long>, dim3(64), dim3(256), 0, nullptr, G, NPAR);
  hipLaunchKernelGGL(zero_kernel<float>, dim3(1), dim3(64), 0, nullptr, d_loss, 1);
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks_for(B)), dim3(64 * kWaves), 0, nullptr,
                     d_params, d_X, d_T, B, 1.f / (float)B, G, d_loss);
  hipLaunchKernelGGL(grad_out_kernel, dim3(64), dim3(256), 0, nullptr, G, d_grads, NPAR);
  float h = 0.f;
  const bool ok = hipMemcpy(&h, d_loss, sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
  (void)hipFree(d_loss);
  (void)hipFree(G);
  if (loss) *loss = h;
  return ok ? PSH_OK : PSH_ERR_DEVICE;
}

extern "C" int psh_mlp_train(float *d_params, const float *d_X, const float *d_T, int B,
                             int batch, int epochs, float lr, unsigned seed, float *loss_out) {
  if (!d_params || !d_X || !d_T || B < 1 || batch < 16 || epochs < 0) return PSH_ERR_ARG;
  float *buf = nullptr;
  long long *G = nullptr;
  if (hipMalloc(&buf, sizeof(float) * (2 * NPAR + 1)) != hipSuccess) return PSH_ERR_MEMORY;
  if (hipMalloc(&G, sizeof(long long) * NPAR) != hipSuccess) {
    (void)hipFree(buf);
    return PSH_ERR_MEMORY;
  }
  float *m = buf, *v = buf + NPAR, *d_loss = buf + 2 * NPAR;
  hipLaunchKernelGGL(zero_kernel<float>, dim3(64), dim3(256), 0, nullptr, buf, 2 * NPAR + 1);
  hipLaunchKernelGGL(zero_kernel<long long>, dim3(64), dim3(256), 0, nullptr, G, NPAR);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  int step = 0;
  // mini-batches are contiguous slices of a sample order the caller shuffles
  // (seed reserved for a device-side shuffle)
  (void)seed;
  for (int ep = 0; ep < epochs; ep++) {
    for (int s0 = 0; s0 < B; s0 += batch) {
      const int nb = B - s0 < batch ? B - s0 : batch;
      hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks_for(nb)), dim3(64 * kWaves), 0, nullptr,
                         d_params, d_X + (size_t)s0 * IN, d_T + s0, nb, 1.f / (float)nb, G,
                         d_loss);
      step++;
      const float c1 = 1.f / (1.f - powf(b1, (float)step));
      const float c2 = 1.f / (1.f - powf(b2, (float)step));
      hipLaunchKernelGGL(adam_kernel, dim3(64), dim3(256), 0, nullptr, d_params, G, m, v, NPAR,
                         lr, b1, b2, eps, c1, c2);
    }
  }
  // final loss over the whole set
  hipLaunchKernelGGL(zero_kernel<float>, dim3(1), dim3(64), 0, nullptr, d_loss, 1);
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks_for(B)), dim3(64 * kWaves), 0, nullptr,
                     d_params, d_X, d_T, B, 1.f / (float)B, G, d_loss);
  float h = 0.f;
  const bool ok = hipMemcpy(&h, d_loss, sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
  (void)hipFree(buf);
  (void)hipFree(G);
  if (loss_out) *loss_out = h;
  return ok ? PSH_OK : PSH_ERR_DEVICE;
}
