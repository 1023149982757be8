// fft_core.h -- register/LDS Stockham FFT building blocks for gfx950.
//
// A workgroup of T threads transforms C lines of N complex points.  Each
// Stockham pass takes radix-R butterflies (R in 2,4,8,16) on elements
// j + r*N/R, multiplies by the twiddles exp(-2 pi i r (j mod Ns) / (Ns R)),
// and writes element (j / Ns) * Ns * R + (j mod Ns) + r * Ns; after the last
// pass the line is in natural order (Govindaraju et al. 2008).  Between passes
// the values are exchanged through LDS one component at a time (real parts,
// then imaginary parts), which halves the LDS footprint of a tile.
//
// Lane mapping: COLMAP (strided columns in HBM) puts consecutive lanes on
// consecutive lines, so that 16 lanes read one 128-byte row segment; ROWMAP
// (contiguous rows) puts consecutive lanes on consecutive elements of a line.
#pragma once

#include <hip/hip_runtime.h>

namespace pshfft {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <bool INV>
__device__ __forceinline__ float2 twmul(float2 a, float2 w) {
  return INV ? cmulc(a, w) : cmul(a, w);
}
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }
constexpr bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
constexpr int pow2_floor(int n) { return n < 2 ? 1 : 2 * pow2_floor(n / 2); }
constexpr int pow2_ceil(int n) { return pow2_floor(n) < n ? 2 * pow2_floor(n) : pow2_floor(n); }
// Lengths N = f 2^a with f in {1, 3, 5}: the power-of-two part and f.
constexpr int pow2_part(int N) { return N & -N; }
constexpr int odd_part(int N) { return N / pow2_part(N); }
// Radix plans of a length-N transform (PL), for the power-of-two part:
//   0  the fewest passes (radix <= 16), the larger radices first;
//   1  the same passes, the smaller radices first;
//   2  one pass more (smaller radices), the larger first;
// an odd factor f = 3 or 5 adds one radix-f pass at the end (its stride
// N / f is then a power of two, which keeps the padded LDS addresses of
// the other passes separable).
constexpr int npasses2(int P, int PL) {
  return (ilog2(P) + 3) / 4 + (PL == 2 && (ilog2(P) + 3) / 4 < ilog2(P) ? 1 : 0);
}
constexpr int npasses(int N, int PL = 0) {
  return npasses2(pow2_part(N), PL) + (odd_part(N) > 1 ? 1 : 0);
}
constexpr int radix_bits(int N, int p, int PL = 0) {
  return ilog2(pow2_part(N)) / npasses2(pow2_part(N), PL) +
         ((PL == 1 ? npasses2(pow2_part(N), PL) - 1 - p : p) <
                  ilog2(pow2_part(N)) % npasses2(pow2_part(N), PL)
              ? 1
              : 0);
}
constexpr int radix(int N, int p, int PL = 0) {
  return (odd_part(N) > 1 && p == npasses(N, PL) - 1) ? odd_part(N)
                                                       : 1 << radix_bits(N, p, PL);
}
constexpr int ns_before(int N, int p, int PL = 0) {
  return p == 0 ? 1 : ns_before(N, p - 1, PL) * radix(N, p - 1, PL);
}

// M - k modulo M (the partner of element k of an M-point line).
template <int M>
__device__ __forceinline__ int mirror(int k) {
  if constexpr (is_pow2(M)) return (M - k) & (M - 1);
  else return k == 0 ? 0 : M - k;
}

// exp(-2 pi i m / 16) for a compile-time m (after unrolling)
__device__ __forceinline__ float2 w16(int m) {
  constexpr float c1 = 0.92387953251128674f, s1 = 0.38268343236508978f;
  constexpr float h = 0.70710678118654752f;
  switch (m & 15) {
    case 0: return make_float2(1.f, 0.f);
    case 1: return make_float2(c1, -s1);
    case 2: return make_float2(h, -h);
    case 3: return make_float2(s1, -c1);
    case 4: return make_float2(0.f, -1.f);
    case 5: return make_float2(-s1, -c1);
    case 6: return make_float2(-h, -h);
    case 7: return make_float2(-c1, -s1);
    case 8: return make_float2(-1.f, 0.f);
    case 9: return make_float2(-c1, s1);
    case 10: return make_float2(-h, h);
    case 11: return make_float2(-s1, c1);
    case 12: return make_float2(0.f, 1.f);
    case 13: return make_float2(s1, c1);
    case 14: return make_float2(h, h);
    default: return make_float2(c1, s1);
  }
}

template <bool INV>
__device__ __forceinline__ void dft2(float2 &a, float2 &b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2 &x0, float2 &x1, float2 &x2,
                                     float2 &x3) {
  float2 a = cadd(x0, x2), b = csub(x0, x2);
  float2 c = cadd(x1, x3), d = rot<INV>(csub(x1, x3));
  x0 = cadd(a, c);
  x2 = csub(a, c);
  x1 = cadd(b, d);
  x3 = csub(b, d);
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2 *v);

template <>
__device__ __forceinline__ void dft<2, false>(float2 *v) { dft2<false>(v[0], v[1]); }
template <>
__device__ __forceinline__ void dft<2, true>(float2 *v) { dft2<true>(v[0], v[1]); }
template <>
__device__ __forceinline__ void dft<4, false>(float2 *v) {
  dft4<false>(v[0], v[1], v[2], v[3]);
}
template <>
__device__ __forceinline__ void dft<4, true>(float2 *v) {
  dft4<true>(v[0], v[1], v[2], v[3]);
}

// 8 = 4 x 2: X[k1 + 4 k2] = sum_n2 W8^(n2 k1) W2^(n2 k2) DFT4(x[2 n1 + n2])
template <bool INV>
__device__ __forceinline__ void dft8(float2 *v) {
  float2 a0 = v[0], a1 = v[2], a2 = v[4], a3 = v[6];
  float2 b0 = v[1], b1 = v[3], b2 = v[5], b3 = v[7];
  dft4<INV>(a0, a1, a2, a3);
  dft4<INV>(b0, b1, b2, b3);
  b1 = twmul<INV>(b1, w16(2));
  b2 = rot<INV>(b2);
  b3 = twmul<INV>(b3, w16(6));
  v[0] = cadd(a0, b0);
  v[4] = csub(a0, b0);
  v[1] = cadd(a1, b1);
  v[5] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[6] = csub(a2, b2);
  v[3] = cadd(a3, b3);
  v[7] = csub(a3, b3);
}
template <>
__device__ __forceinline__ void dft<8, false>(float2 *v) { dft8<false>(v); }
template <>
__device__ __forceinline__ void dft<8, true>(float2 *v) { dft8<true>(v); }

// 3-point DFT: X1,2 = x0 - (x1 + x2)/2 -+ i (sqrt 3 / 2)(x1 - x2) (forward)
template <bool INV>
__device__ __forceinline__ void dft3(float2 *v) {
  constexpr float s3 = 0.86602540378443865f;
  const float2 t1 = cadd(v[1], v[2]);
  const float2 t2 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
  const float2 t3 = rot<INV>(csub(v[1], v[2]));
  v[0] = cadd(v[0], t1);
  v[1] = make_float2(t2.x + s3 * t3.x, t2.y + s3 * t3.y);
  v[2] = make_float2(t2.x - s3 * t3.x, t2.y - s3 * t3.y);
}
template <>
__device__ __forceinline__ void dft<3, false>(float2 *v) { dft3<false>(v); }
template <>
__device__ __forceinline__ void dft<3, true>(float2 *v) { dft3<true>(v); }

// 5-point DFT: with c_k = cos(2 pi k / 5), s_k = sin(2 pi k / 5),
// t1 = x1 + x4, t2 = x2 + x3, t3 = x1 - x4, t4 = x2 - x3:
// X1,4 = x0 + c1 t1 + c2 t2 -+ i (s1 t3 + s2 t4),
// X2,3 = x0 + c2 t1 + c1 t2 -+ i (s2 t3 - s1 t4) (forward).
template <bool INV>
__device__ __forceinline__ void dft5(float2 *v) {
  constexpr float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
  constexpr float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
  const float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  const float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  const float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
  const float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
  const float2 b1 = rot<INV>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
  const float2 b2 = rot<INV>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
  v[0] = make_float2(v[0].x + t1.x + t2.x, v[0].y + t1.y + t2.y);
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}
template <>
__device__ __forceinline__ void dft<5, false>(float2 *v) { dft5<false>(v); }
template <>
__device__ __forceinline__ void dft<5, true>(float2 *v) { dft5<true>(v); }

// 16 = 4 x 4
template <bool INV>
__device__ __forceinline__ void dft16(float2 *v) {
  float2 y[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; n2++) {
    y[n2][0] = v[n2];
    y[n2][1] = v[4 + n2];
    y[n2][2] = v[8 + n2];
    y[n2][3] = v[12 + n2];
    dft4<INV>(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
  }
#pragma unroll
  for (int n2 = 1; n2 < 4; n2++)
#pragma unroll
    for (int k1 = 1; k1 < 4; k1++) y[n2][k1] = twmul<INV>(y[n2][k1], w16(n2 * k1));
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) {
    float2 a = y[0][k1], b = y[1][k1], c = y[2][k1], d = y[3][k1];
    dft4<INV>(a, b, c, d);
    v[k1] = a;
    v[k1 + 4] = b;
    v[k1 + 8] = c;
    v[k1 + 12] = d;
  }
}
template <>
__device__ __forceinline__ void dft<16, false>(float2 *v) { dft16<false>(v); }
template <>
__device__ __forceinline__ void dft<16, true>(float2 *v) { dft16<true>(v); }

// LDS index of element n of line c (one float per element, padded).
// COLMAP (lanes vary the line fastest): element-major [n][c] with 8 floats of
// padding every 8 elements; ROWMAP (lanes vary the element fastest):
// line-major [c][n] with one float of padding every 16 elements and an odd
// line stride.  Both were chosen with a bank-conflict simulation of every
// Stockham exchange of the plans used here (32 and 64 banks).
template <int N, int C, bool COLMAP>
struct Lds {
  static constexpr int LN = N + (N >> 4) + 1;
  static constexpr int SIZE = COLMAP ? N * C + ((N + 7) >> 3) * 8 : C * LN;
  __device__ static __forceinline__ int idx(int c, int n) {
    return COLMAP ? n * C + c + ((n >> 3) << 3) : c * LN + n + (n >> 4);
  }
  // idx(c, n0 + d) - idx(c, n0) for the element groups the passes touch:
  // a butterfly's elements n0 + r*S (n0 mod S < S, S and the radix powers of
  // two) never carry across a padding block (16 rows, or 8 columns), so the
  // difference is this constant and an element's address is one per-thread
  // base plus an immediate offset.
  __device__ static constexpr int off(int d) {
    return COLMAP ? d * C + ((d >> 3) << 3) : d + (d >> 4);
  }
};

// threadIdx.x hidden from loop-invariant code motion.  In persistent kernels
// every per-lane address derived from the thread index is invariant over the
// tile loop; hoisted, those addresses stay live across the whole loop and
// spill.  Recomputing them at each use costs a few integer operations.
// The mask restores the range 0 <= t < T (T a power of two, the block size)
// that the bound checks of the geometry below fold against.
template <int T>
__device__ __forceinline__ int opaque_tid() {
  static_assert((T & (T - 1)) == 0, "power-of-two block size");
  int t = (int)threadIdx.x;
  asm volatile("" : "+v"(t));
  return t & (T - 1);
}

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8
// XCDs (linear workgroup id mod 8), each with its own L2; mapping workgroup b
// of nblk to tile (b mod 8) * nblk/8 + b/8 gives every XCD a contiguous range
// of tiles, so neighbouring tiles -- which share 128-byte lines -- meet in one
// L2.  Identity when nblk is not a multiple of 8.
__device__ __forceinline__ size_t xcd_tile(size_t b, size_t nblk) {
  if (nblk & 7) return b;
  return (b & 7) * (nblk >> 3) + (b >> 3);
}

// Per-pass thread geometry.
template <int N, int C, int T, int P, bool COLMAP, int PL = 0>
struct PassGeom {
  static constexpr int R = radix(N, P, PL);
  static constexpr int NS = ns_before(N, P, PL);
  static constexpr int NBF = N / R;           // butterflies per line
  static constexpr int NBT = C * NBF;         // butterflies in the tile
  static constexpr int NB = (NBT + T - 1) / T;// butterflies per thread
  __device__ static __forceinline__ bool map(int q, int &c, int &j) {
    int b = opaque_tid<T>() + q * T;
    if (COLMAP) {
      c = b % C;
      j = b / C;
    } else {
      c = b / NBF;
      j = b % NBF;
    }
    return b < NBT;
  }
};

// Values per thread (same for all passes of a plan).
template <int N, int C, int T>
struct Plan {
  static constexpr int P = npasses(N);
  static constexpr int V = ((C * N + T - 1) / T) > 0 ? (C * N + T - 1) / T : 1;
};

// Largest per-thread value count over the passes of a plan.
template <int N, int C, int T, bool COLMAP, int P = 0, int PL = 0>
constexpr int max_values() {
  if constexpr (P >= npasses(N, PL)) {
    return 1;
  } else {
    constexpr int here =
        PassGeom<N, C, T, P, COLMAP, PL>::NB * PassGeom<N, C, T, P, COLMAP, PL>::R;
    constexpr int rest = max_values<N, C, T, COLMAP, P + 1, PL>();
    return here > rest ? here : rest;
  }
}

// Load the inputs of pass P through `ld(c, n) -> float2`.
template <int N, int C, int T, int P, bool COLMAP, class LD, int PL = 0>
__device__ __forceinline__ void pass_load(float2 *v, LD ld) {
  using G = PassGeom<N, C, T, P, COLMAP, PL>;
#pragma unroll
  for (int q = 0; q < G::NB; q++) {
    int c, j;
    bool ok = G::map(q, c, j);
#pragma unroll
    for (int r = 0; r < G::R; r++)
      v[q * G::R + r] = ok ? ld(c, j + r * G::NBF) : make_float2(0.f, 0.f);
  }
}

// Twiddle and butterfly of pass P in registers.  tw: table of
// exp(-2 pi i m / (N*TWS)) read with stride TWS.
template <int N, int C, int T, int P, bool COLMAP, bool INV, int TWS, int PL = 0>
__device__ __forceinline__ void pass_compute(float2 *v,
                                             const float2 *__restrict__ tw) {
  using G = PassGeom<N, C, T, P, COLMAP, PL>;
#pragma unroll
  for (int q = 0; q < G::NB; q++) {
    int c, j;
    G::map(q, c, j);
    if (P > 0) {
      const int k = (int)((unsigned)j % (unsigned)G::NS);
      constexpr int step = N / (G::NS * G::R);
#pragma unroll
      for (int r = 1; r < G::R; r++)
        v[q * G::R + r] = twmul<INV>(v[q * G::R + r], tw[(r * k * step) * TWS]);
    }
    dft<G::R, INV>(v + q * G::R);
  }
}

// Store the outputs of pass P through `st(c, n, value)`.
template <int N, int C, int T, int P, bool COLMAP, class ST, int PL = 0>
__device__ __forceinline__ void pass_store(const float2 *v, ST st) {
  using G = PassGeom<N, C, T, P, COLMAP, PL>;
#pragma unroll
  for (int q = 0; q < G::NB; q++) {
    int c, j;
    if (!G::map(q, c, j)) continue;
    const int k = (int)((unsigned)j % (unsigned)G::NS);
    const int base = (int)((unsigned)j / (unsigned)G::NS) * G::NS * G::R + k;
#pragma unroll
    for (int r = 0; r < G::R; r++) st(c, base + r * G::NS, v[q * G::R + r]);
  }
}

// Exchange the outputs of pass P into the inputs of pass P+1 through LDS,
// one component at a time.
template <int N, int C, int T, int P, bool COLMAP, int PL = 0>
__device__ __forceinline__ void exchange(float2 *v, float *lds) {
  using G0 = PassGeom<N, C, T, P, COLMAP, PL>;
  using G1 = PassGeom<N, C, T, P + 1, COLMAP, PL>;
  using L = Lds<N, C, COLMAP>;
  constexpr int V = G1::NB * G1::R;
  // padded addresses of a butterfly's elements are one base plus constant
  // offsets when the stride is a power of two (or a multiple of the padding
  // block); otherwise each element's address is formed on its own
  constexpr int BLK = COLMAP ? 8 : 16;
  constexpr bool SEP0 = is_pow2(G0::NS) || G0::NS % BLK == 0;
  constexpr bool SEP1 = is_pow2(G1::NBF) || G1::NBF % BLK == 0;
  float tmp[V];
  // outputs of pass P: element base + r NS of line c
  auto store = [&](bool im) {
#pragma unroll
    for (int q = 0; q < G0::NB; q++) {
      int c, j;
      if (!G0::map(q, c, j)) continue;
      const int k = (int)((unsigned)j % (unsigned)G0::NS);
      const int e0 = (int)((unsigned)j / (unsigned)G0::NS) * G0::NS * G0::R + k;
      float *row = lds + L::idx(c, e0);
#pragma unroll
      for (int r = 0; r < G0::R; r++) {
        const float2 x = v[q * G0::R + r];
        if (SEP0) row[L::off(r * G0::NS)] = im ? x.y : x.x;
        else lds[L::idx(c, e0 + r * G0::NS)] = im ? x.y : x.x;
      }
    }
  };
  // inputs of pass P+1: element j + r NBF of line c
  auto load = [&](bool im) {
#pragma unroll
    for (int q = 0; q < G1::NB; q++) {
      int c, j;
      const bool ok = G1::map(q, c, j);
      const float *row = lds + L::idx(c, j);
#pragma unroll
      for (int r = 0; r < G1::R; r++) {
        const float x = ok ? (SEP1 ? row[L::off(r * G1::NBF)] : lds[L::idx(c, j + r * G1::NBF)])
                           : 0.f;
        if (im) v[q * G1::R + r] = make_float2(tmp[q * G1::R + r], x);
        else tmp[q * G1::R + r] = x;
      }
    }
  };
  store(false);
  __syncthreads();
  load(false);
  __syncthreads();
  store(true);
  __syncthreads();
  load(true);
  __syncthreads();
}

// Run passes P..last on values already loaded for pass P.
template <int N, int C, int T, int P, bool COLMAP, bool INV, int TWS, int PL = 0>
struct Passes {
  __device__ static __forceinline__ void run(float2 *v, float *lds,
                                             const float2 *__restrict__ tw) {
    pass_compute<N, C, T, P, COLMAP, INV, TWS, PL>(v, tw);
    if constexpr (P + 1 < npasses(N, PL)) {
      exchange<N, C, T, P, COLMAP, PL>(v, lds);
      Passes<N, C, T, P + 1, COLMAP, INV, TWS, PL>::run(v, lds, tw);
    }
  }
};

// Full transform: load pass-0 inputs with ld, run all passes, leave the
// outputs of the last pass in v (use pass_store<..., npasses-1> to place them).
template <int N, int C, int T, bool COLMAP, bool INV, int TWS, class LD>
__device__ __forceinline__ void transform(float2 *v, float *lds,
                                          const float2 *__restrict__ tw, LD ld) {
  pass_load<N, C, T, 0, COLMAP>(v, ld);
  Passes<N, C, T, 0, COLMAP, INV, TWS>::run(v, lds, tw);
}

template <int N, int C, int T, bool COLMAP, class ST>
__device__ __forceinline__ void store_last(const float2 *v, ST st) {
  pass_store<N, C, T, npasses(N) - 1, COLMAP>(v, st);
}

// Row-split access for strided columns.  Element j + r*S of line c (S = NBF
// of the first pass for loads, NS of the last pass for stores) is passed as
// (r, c, j): with r a compile-time constant after unrolling, a caller forms
// the address as a per-tile uniform row base (r) plus one per-lane offset
// (c, j) shared by all values of the thread, which keeps 64-bit addresses
// out of the registers.
template <int N, int C, int T, bool COLMAP, class LD, int PL = 0>
__device__ __forceinline__ void pass_load_rows(float2 *v, LD ld) {
  using G = PassGeom<N, C, T, 0, COLMAP, PL>;
#pragma unroll
  for (int q = 0; q < G::NB; q++) {
    int c, j;
    const bool ok = G::map(q, c, j);
#pragma unroll
    for (int r = 0; r < G::R; r++)
      v[q * G::R + r] = ok ? ld(r, c, j) : make_float2(0.f, 0.f);
  }
}

template <int N, int C, int T, bool COLMAP, bool INV, int TWS, class LD, int PL = 0>
__device__ __forceinline__ void transform_rows(float2 *v, float *lds,
                                               const float2 *__restrict__ tw, LD ld) {
  pass_load_rows<N, C, T, COLMAP, LD, PL>(v, ld);
  Passes<N, C, T, 0, COLMAP, INV, TWS, PL>::run(v, lds, tw);
}

// Two fields of the same tile advanced pass by pass together (each exchanged
// through the same LDS buffer in turn), so that neither sits finished in
// registers while the other is transformed.
template <int N, int C, int T, int P, bool COLMAP, bool INV, int TWS>
struct Passes2 {
  __device__ static __forceinline__ void run(float2 *va, float2 *vb, float *lds,
                                             const float2 *__restrict__ tw) {
    pass_compute<N, C, T, P, COLMAP, INV, TWS>(va, tw);
    pass_compute<N, C, T, P, COLMAP, INV, TWS>(vb, tw);
    if constexpr (P + 1 < npasses(N)) {
      exchange<N, C, T, P, COLMAP>(va, lds);
      exchange<N, C, T, P, COLMAP>(vb, lds);
      Passes2<N, C, T, P + 1, COLMAP, INV, TWS>::run(va, vb, lds, tw);
    }
  }
};

template <int N, int C, int T, bool COLMAP, bool INV, int TWS, class LDA, class LDB>
__device__ __forceinline__ void transform_rows2(float2 *va, float2 *vb, float *lds,
                                                const float2 *__restrict__ tw, LDA lda,
                                                LDB ldb) {
  pass_load_rows<N, C, T, COLMAP>(va, lda);
  pass_load_rows<N, C, T, COLMAP>(vb, ldb);
  Passes2<N, C, T, 0, COLMAP, INV, TWS>::run(va, vb, lds, tw);
}

// Natural-order outputs of the last pass: element j + r*NS of line c.
template <int N, int C, int T, bool COLMAP, class ST, int PL = 0>
__device__ __forceinline__ void store_last_rows(const float2 *v, ST st) {
  using G = PassGeom<N, C, T, npasses(N, PL) - 1, COLMAP, PL>;
  static_assert(G::NS * G::R == N, "last pass");
#pragma unroll
  for (int q = 0; q < G::NB; q++) {
    int c, j;
    if (!G::map(q, c, j)) continue;
#pragma unroll
    for (int r = 0; r < G::R; r++) st(r, c, j, v[q * G::R + r]);
  }
}

// Access to element `off` of a wave-uniform row base.  When every lane
// offset fits 32 bits in bytes (BYTE32), the address is formed as a 64-bit
// scalar base plus a zero-extended 32-bit byte offset, which the global
// load/store instructions take directly (SGPR base + VGPR offset): one VGPR
// per lane instead of a 64-bit address per access.
template <bool BYTE32>
__device__ __forceinline__ const float2 *row_elem(const float2 *rowbase, uint32_t off) {
  if constexpr (BYTE32)
    return reinterpret_cast<const float2 *>(reinterpret_cast<const char *>(rowbase) +
                                            (size_t)(off * 8u));
  else
    return rowbase + off;
}
template <bool BYTE32>
__device__ __forceinline__ float2 *row_elem(float2 *rowbase, uint32_t off) {
  return const_cast<float2 *>(row_elem<BYTE32>(static_cast<const float2 *>(rowbase), off));
}

// Mesh traffic of the column, z and x passes.  A launch whose meshes
// outgrow the 256 MB Infinity Cache reads and writes every element once, so
// its loads and stores are non-temporal (the `nt` bit): measured on MI355X
// at N = 1024 the y pass drops from 4.1 to 3.4 ms and the whole path by 4 %.
// From N = 512 one mesh (537 MB per grid) is that large; smaller meshes
// stream when a batch of them is.  Meshes that fit stay cached between
// passes and keep ordinary accesses (non-temporal ones cost 2 % for one
// N = 256 pair).
template <int N>
constexpr bool stream_mesh() { return N >= 512; }

constexpr size_t kStreamBytes = (size_t)256 << 20;

inline bool stream_bytes(size_t bytes) { return bytes > kStreamBytes; }

typedef float v2f_t __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ __forceinline__ float2 mesh_load(const float2 *p) {
  if constexpr (NT) {
    const v2f_t v = __builtin_nontemporal_load(reinterpret_cast<const v2f_t *>(p));
    return make_float2(v.x, v.y);
  } else {
    return *p;
  }
}

template <bool NT>
__device__ __forceinline__ void mesh_store(float2 *p, float2 x) {
  if constexpr (NT) {
    const v2f_t v = {x.x, x.y};
    __builtin_nontemporal_store(v, reinterpret_cast<v2f_t *>(p));
  } else {
    *p = x;
  }
}

// Non-temporal 16-byte store (two complex values).
__device__ __forceinline__ void mesh_store4(float4 *p, float4 x) {
  typedef float v4f_t __attribute__((ext_vector_type(4)));
  const v4f_t v = {x.x, x.y, x.z, x.w};
  __builtin_nontemporal_store(v, reinterpret_cast<v4f_t *>(p));
}

// Row strides of the split access.
template <int N, int C, int T, bool COLMAP, int PL = 0>
struct RowSplit {
  static constexpr int LOAD = PassGeom<N, C, T, 0, COLMAP, PL>::NBF;
  static constexpr int STORE = PassGeom<N, C, T, npasses(N, PL) - 1, COLMAP, PL>::NS;
  // largest per-lane row index (must keep j * ES + c within 32 bits)
  static constexpr int JMAX = (LOAD > STORE ? LOAD : STORE);
  // lane offsets of a column pass over an x stride of N (N/2 + pad) complex
  // values fit 32 bits in elements always, in bytes when BYTE32
  static constexpr size_t XSPAN = (size_t)JMAX * N * (N / 2 + 16) + C;
  static constexpr bool BYTE32 = XSPAN * 8 < (1ull << 32);
};

}  // namespace pshfft
