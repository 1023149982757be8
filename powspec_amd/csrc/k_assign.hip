// k_assign.hip -- NGP/CIC/TSC/PCS mass assignment with grid interlacing on
// gfx950, replacing the reference's per-particle `omp atomic` scatter
// (src/genr_mesh.c:50-412) and its zero/subtract/shift passes
// (src/genr_mesh.c:590-602, 793-858).
//
// Design:
//  1. count:   per particle, the FP64 cell coordinate decides the base cell of
//              the window on the plain grid and on the half-cell-shifted grid
//              (same operation order as the reference, contraction off); the
//              union window is mapped to the 1..8 mesh tiles it touches and a
//              per-tile counter is incremented.
//  2. scan:    exclusive prefix sum of the tile counters.
//  3. fill:    each (particle, tile) pair is written as a 24-byte record.
//  4. deposit: one workgroup per tile accumulates its records into an LDS
//              copy of the tile (fixed-point integer LDS atomics, both grids)
//              and writes the tile to HBM once with coalesced row stores.
//              Every mesh cell is written exactly once: no zeroing pass, no
//              global atomics.
// Surveys assign data and randoms into the same tiles with weights w and
// -alpha*w, which fuses the data-minus-random subtraction
// (src/genr_mesh.c:802-806, 841-845).
#include "psh_internal.h"
#include "fft_core.h"

#include <type_traits>

namespace {

struct __align__(8) Rec {
  float f[3];      // window fraction on the plain grid, per axis
  float w;         // particle weight times the catalogue scale
  uint32_t bxy;    // base cell x | y << 16 (plain grid)
  uint32_t bz;     // base cell z | delta bits (shifted grid) << 16
};
static_assert(sizeof(Rec) == 24, "record layout");

using DevPart = PshDevPart;

__device__ __forceinline__ int window_width(int scheme) { return scheme + 1; }

// Periodic wrap into [0, Ng).  Bases and window ends are at most a window
// (or a grid) away from [0, Ng), so one compare settles nearly every call;
// the remainder covers coordinates further out.
__device__ __forceinline__ int wrap(int c, int Ng) {
  if ((unsigned)c < (unsigned)Ng) return c;
  const int r = c % Ng;
  return r < 0 ? r + Ng : r;
}

// Base cell (leftmost window cell, unwrapped) and window fraction of a grid
// coordinate x >= 0, following src/genr_mesh.c:57-66 (NGP), 94-104 (CIC),
// 152-172 (TSC), 256-263 (PCS).  The fraction is d = x - (int)x for CIC/PCS
// and the signed offset from the middle node for TSC.
__device__ __forceinline__ void axis_base(double x, int scheme, int &b,
                                          double &f) {
  int x0 = (int)x;
  double d = x - x0;
  switch (scheme) {
    case 0:
      b = (d >= 0.5) ? x0 + 1 : x0;
      f = 0;
      break;
    case 1:
      b = x0;
      f = d;
      break;
    case 2:
      if (d < 0.5) {
        b = x0 - 1;
        f = d;
      } else {
        b = x0;
        f = d - 1;
      }
      break;
    default:
      b = x0 - 1;
      f = d;
      break;
  }
}

// Per-particle plain/shifted bases.  The shifted grid uses the reference's
// periodic shift (src/genr_mesh.c:595-600: x >= smin + L -> x - L) before
// the coordinate (x - smin) * Ng / L.
struct PartCells {
  int b[3];        // plain-grid base, wrapped into [0, Ng)
  int delta[3];    // shifted base - plain base (mod Ng), 0 or 1
  float f[3];
  bool ok;
};

__device__ PartCells part_cells(const double *r, const PshAssignGeom &g) {
#pragma clang fp contract(off)
  PartCells pc;
  pc.ok = true;
  const double Ngd = (double)g.Ng;
  for (int a = 0; a < 3; a++) {
    double x = (r[a] - g.bmin[a]) * Ngd / g.L[a];
    int b;
    double f;
    axis_base(x, g.scheme, b, f);
    pc.b[a] = wrap(b, g.Ng);
    pc.f[a] = (float)f;
    pc.delta[a] = 0;
    if (g.intlace) {
      double p = r[a];
      if (p >= g.smin[a] + g.L[a]) p -= g.L[a];
      double xs = (p - g.smin[a]) * Ngd / g.L[a];
      int bs;
      double fs;
      axis_base(xs, g.scheme, bs, fs);
      int dlt = wrap(bs - b, g.Ng);
      if (dlt > 1) pc.ok = false;
      pc.delta[a] = dlt;
    }
  }
  return pc;
}

// Tiles touched along one axis by the union window [b, b + W - 1 + delta]
// (b in [0, Ng), tiles of 2^TS cells).
__device__ __forceinline__ int axis_tiles(int b, int W, int delta, int Ng,
                                          int TS, int t[2]) {
  int e = b + W - 1 + delta;
  t[0] = b >> TS;
  int te = wrap(e, Ng) >> TS;
  if (te != t[0]) {
    t[1] = te;
    return 2;
  }
  return 1;
}

__device__ __forceinline__ int find_part(const DevPart *parts, int nparts,
                                         size_t gi) {
  int lo = 0, hi = nparts - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (parts[mid].start <= gi) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// One particle as loaded: position and weight, its part's weight scale and
// destination mock.
struct Loaded {
  double r[4];
  float wscale;
  int mock;
};

// A part descriptor held in scalar registers (the part of a workgroup's
// first particle).
struct UPart {
  const double *xyzw;
  size_t start, n;
  float wscale;
  int mock;
};

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// The part containing particle `gi0`; every lane passes the same gi0.
__device__ __forceinline__ UPart uniform_part(const DevPart *parts, int nparts, size_t gi0) {
  const DevPart &P = parts[find_part(parts, nparts, gi0)];
  UPart u;
  u.xyzw = reinterpret_cast<const double *>(uniform64(reinterpret_cast<uint64_t>(P.xyzw)));
  u.start = uniform64(P.start);
  u.n = uniform64(P.n);
  u.wscale = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(P.wscale)));
  u.mock = __builtin_amdgcn_readfirstlane(P.mock);
  return u;
}

typedef __attribute__((address_space(1))) const double gdouble;

// Load particle gi (zeros when !live).  P0 is the part the caller expects gi
// to lie in (chunks rarely cross part boundaries), so the table search runs
// only for particles outside it; callers keep the next particle's loads in
// flight while they process the current one.
__device__ __forceinline__ Loaded load_particle(const DevPart *parts, int nparts, UPart P0,
                                                size_t gi, bool live) {
  Loaded L;
  L.r[0] = L.r[1] = L.r[2] = L.r[3] = 0.0;
  L.wscale = 0.f;
  L.mock = 0;
  if (!live) return L;
  const double *px = P0.xyzw;
  size_t ps = P0.start;
  float pw = P0.wscale;
  int pm = P0.mock;
  if (gi < P0.start || gi - P0.start >= P0.n) {
    const int q = find_part(parts, nparts, gi);
    px = parts[q].xyzw;
    ps = parts[q].start;
    pw = parts[q].wscale;
    pm = parts[q].mock;
  }
  L.wscale = pw;
  L.mock = pm;
  const gdouble *g = (const gdouble *)(px + 4 * (gi - ps));   // device memory
  L.r[0] = g[0];
  L.r[1] = g[1];
  L.r[2] = g[2];
  L.r[3] = g[3];
  return L;
}

// Tiles (mock-local indices) touched by a particle rr = (x, y, z, w) and,
// with FILL, its record.  Returns the number of tiles (0 and flag bit 2 when the
// shifted grid's base is inconsistent).  With CHECK and g.check_box, a
// coordinate outside [bmin, bmin + L) (bmin = 0 for simulation boxes) sets
// flag bit 4: the box check of
// simulation catalogues (src/genr_mesh.c:517-533) folded into the count
// pass, which reads every particle anyway (NaN passes, as it passes the
// reference's min/max comparisons).
template <bool FILL, bool CHECK = false>
__device__ __forceinline__ int particle_tiles(const PshAssignGeom &g, const double *rr,
                                              float wscale, int W, uint32_t *tiles, Rec &rec,
                                              int *flags) {
  if (CHECK && g.check_box) {
    bool bad = false;
#pragma unroll
    for (int a = 0; a < 3; a++) bad |= (rr[a] < g.bmin[a]) || (rr[a] >= g.bmin[a] + g.L[a]);
    if (bad) atomicOr(flags, 4);
  }
  PartCells pc = part_cells(rr, g);
  if (!pc.ok) {
    atomicOr(flags, 2);
    return 0;
  }
  int tx[2], ty[2], tz[2];
  int nx = axis_tiles(pc.b[0], W, pc.delta[0], g.Ng, g.tile_shift[0], tx);
  int ny = axis_tiles(pc.b[1], W, pc.delta[1], g.Ng, g.tile_shift[1], ty);
  int nz = axis_tiles(pc.b[2], W, pc.delta[2], g.Ng, g.tile_shift[2], tz);
  int nt = 0;
  for (int a = 0; a < nx; a++)
    for (int b = 0; b < ny; b++)
      for (int c = 0; c < nz; c++)
        tiles[nt++] = ((uint32_t)tx[a] * g.ntile[1] + ty[b]) * g.ntile[2] + tz[c];
  if (FILL) {
    rec.f[0] = pc.f[0];
    rec.f[1] = pc.f[1];
    rec.f[2] = pc.f[2];
    rec.w = (float)rr[3] * wscale;
    rec.bxy = (uint32_t)pc.b[0] | ((uint32_t)pc.b[1] << 16);
    rec.bz = (uint32_t)pc.b[2] |
             ((uint32_t)(pc.delta[0] | (pc.delta[1] << 1) | (pc.delta[2] << 2)) << 16);
  }
  return nt;
}

// Largest |weight| of a launch, for the fixed-point scale of the deposit
// (flags[1] holds the bits of a non-negative float).
__device__ __forceinline__ void publish_wmax(float wmax, int *flags) {
  for (int o = 32; o > 0; o >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, o));
  if (__lane_id() == 0) atomicMax(reinterpret_cast<unsigned *>(flags) + 1, __float_as_uint(wmax));
}

// Run-aggregated increment of LDS or global counters: lanes with `act` whose counter index `t`
// equals the previous lane's form a run (space-ordered catalogues put
// neighbouring particles into one tile, so a wave holds a few long runs);
// the run's first lane adds the run length with one LDS atomic and every
// lane gets old + its rank in the run (RET), or nothing (count only).
// Shuffled catalogues make every lane its own run: one atomic per lane, as
// without aggregation.
template <bool RET>
__device__ __forceinline__ uint32_t run_increment(uint32_t *hist, uint32_t t, bool act) {
  const int lane = __lane_id();
  const uint32_t tp = __shfl_up(t, 1, 64);
  const bool actp = __shfl_up(act ? 1 : 0, 1, 64) != 0;
  const bool head = act && (lane == 0 || !actp || tp != t);
  const uint64_t heads = __ballot(head);
  const uint64_t bound = heads | ~__ballot(act);   // a run ends before these lanes
  uint32_t old = 0;
  if (head) {
    const uint64_t after = lane == 63 ? 0ull : bound & (~0ull << (lane + 1));
    const int next = after ? __ffsll((unsigned long long)after) - 1 : 64;
    const uint32_t len = (uint32_t)(next - lane);
    if (RET) old = atomicAdd(&hist[t], len);
    else atomicAdd(&hist[t], len);
  }
  if (!RET) return 0u;
  // first lane of this lane's run: the highest head at or below it
  const uint64_t below = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
  const int hl = below ? 63 - __clzll((long long)below) : lane;
  const uint32_t base = __shfl(old, hl, 64);
  return base + (uint32_t)(lane - hl);
}

template <bool FILL>
__global__ void __launch_bounds__(256)
assign_bucket(PshAssignGeom g, const DevPart *__restrict__ parts, int nparts,
              size_t ntotal, uint32_t *__restrict__ counts,
              uint32_t *__restrict__ cursors, Rec *__restrict__ recs,
              size_t rec_cap, int *__restrict__ flags) {
  const int W = window_width(g.scheme);
  const uint32_t tiles_per_mock = (uint32_t)g.ntile[0] * g.ntile[1] * g.ntile[2];
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  // every lane of a wave runs the same number of iterations so that the
  // wave-level aggregation below sees converged waves
  const size_t iters = (ntotal + stride - 1) / stride;
  const UPart P0 = uniform_part(parts, nparts, blockIdx.x * (size_t)blockDim.x);
  size_t gi = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  float wmax = 0.f;
  // software pipeline: the next particle's loads are in flight while the
  // current one is processed
  bool cur_live = gi < ntotal;
  Loaded cur = load_particle(parts, nparts, P0, gi, cur_live);
#pragma unroll 1
  for (size_t it = 0; it < iters; it++, gi += stride) {
    const size_t gn = gi + stride;
    const bool nxt_live = it + 1 < iters && gn < ntotal;
    const Loaded nxt = load_particle(parts, nparts, P0, gn, nxt_live);
    uint32_t tiles[8];
    int nt = 0;
    Rec rec;
    if (cur_live) {
      nt = particle_tiles<FILL, !FILL>(g, cur.r, cur.wscale, W, tiles, rec, flags);
      const uint32_t mbase = (uint32_t)cur.mock * tiles_per_mock;
      for (int k = 0; k < nt; k++) tiles[k] += mbase;
      if (FILL && nt) wmax = fmaxf(wmax, fabsf(rec.w));
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const bool act = k < nt;
      if (!__ballot(act)) break;
      const uint32_t t = act ? tiles[k] : 0u;
      if (FILL) {
        const uint32_t slot = run_increment<true>(cursors, t, act);
        if (act) {
          if (slot < rec_cap) recs[slot] = rec;
          else atomicOr(flags, 1);
        }
      } else {
        run_increment<false>(counts, t, act);
      }
    }
    cur = nxt;
    cur_live = nxt_live;
  }
  if (FILL) publish_wmax(wmax, flags);
}

// Bucketing with an LDS histogram of one mock's tiles (meshes of at most
// kLdsTiles tiles per mock).  Workgroup b takes particles [b kChunk,
// (b+1) kChunk) -- a split that depends on the total only, so captured
// graphs stay valid for any mock sizes -- and histograms the (particle,
// tile) pairs of the mock of its first particle in LDS; particles of another
// mock (at a part boundary) use global atomics directly.
//  count: the histogram is added to the global counters, one atomic per
//         nonzero tile;
//  fill:  one atomic per nonzero tile reserves the workgroup's range of each
//         bucket, then a second sweep over the chunk (L2-resident) hands out
//         slots with LDS atomics.
// The deposit accumulates in integers, so the order of records within a
// bucket does not matter.
constexpr int kChunk = 16384;
constexpr int kLdsTiles = 32768;
constexpr int kLdsThreads = 1024;

template <bool FILL>
__global__ void __launch_bounds__(kLdsThreads, 8)   // 8 waves per SIMD: two workgroups per CU
assign_bucket_lds(PshAssignGeom g, const DevPart *__restrict__ parts, int nparts,
                  size_t ntotal, uint32_t *__restrict__ counts,
                  uint32_t *__restrict__ cursors, Rec *__restrict__ recs, size_t rec_cap,
                  int *__restrict__ flags, uint32_t *__restrict__ whist) {
  extern __shared__ uint32_t hist[];
  const int W = window_width(g.scheme);
  const uint32_t tpm = (uint32_t)g.ntile[0] * g.ntile[1] * g.ntile[2];
  const size_t c0 = (size_t)blockIdx.x * kChunk;
  const size_t c1 = min(ntotal, c0 + (size_t)kChunk);
  const UPart P0 = uniform_part(parts, nparts, c0);
  const int mock0 = P0.mock;
  const uint32_t mbase0 = (uint32_t)mock0 * tpm;
  uint32_t *wh = whist ? whist + (size_t)blockIdx.x * tpm : nullptr;
  uint32_t tiles[8];
  Rec rec;
  if (FILL && wh) {
    // the count pass left this workgroup's histogram in HBM: reserve the
    // workgroup's range of each bucket without another pass over the chunk
    for (uint32_t i = threadIdx.x; i < tpm; i += blockDim.x) {
      const uint32_t v = wh[i];
      hist[i] = v ? atomicAdd(&cursors[mbase0 + i], v) : 0u;
    }
    __syncthreads();
  } else {
    for (uint32_t i = threadIdx.x; i < tpm; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    // every lane of a wave runs the same number of iterations (the chunk is a
    // multiple of the workgroup size), so the run aggregation sees whole waves
    static_assert(kChunk % kLdsThreads == 0, "chunk of whole workgroup sweeps");
    // pass A: histogram of this mock's pairs (count: others go to global)
    {
      // software pipeline: the next particle's loads are in flight while the
      // current one is processed
      size_t gi = c0 + threadIdx.x;
      bool cur_live = gi < c1;
      Loaded cur = load_particle(parts, nparts, P0, gi, cur_live);
#pragma unroll 1
      for (int it = 0; it < kChunk / kLdsThreads; it++) {
        const size_t gn = gi + kLdsThreads;
        const bool nxt_live = it + 1 < kChunk / kLdsThreads && gn < c1;
        const Loaded nxt = load_particle(parts, nparts, P0, gn, nxt_live);
        int nt = 0;
        int mock = mock0;
        if (cur_live) {
          nt = particle_tiles<false, !FILL>(g, cur.r, cur.wscale, W, tiles, rec, flags);
          mock = cur.mock;
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const bool act = k < nt;
          if (!__ballot(act)) break;
          const bool here = act && mock == mock0;
          run_increment<false>(hist, here ? tiles[k] : 0u, here);
          if (!FILL && act && !here) atomicAdd(&counts[(uint32_t)mock * tpm + tiles[k]], 1u);
        }
        cur = nxt;
        cur_live = nxt_live;
        gi = gn;
      }
    }
    __syncthreads();
    if (!FILL) {
      for (uint32_t i = threadIdx.x; i < tpm; i += blockDim.x) {
        const uint32_t v = hist[i];
        if (wh) wh[i] = v;
        if (v) atomicAdd(&counts[mbase0 + i], v);
      }
      return;
    }
    for (uint32_t i = threadIdx.x; i < tpm; i += blockDim.x)
      if (hist[i]) hist[i] = atomicAdd(&cursors[mbase0 + i], hist[i]);
    __syncthreads();
  }   // without a stored histogram
  // pass B: records
  float wmax = 0.f;
  {
    // software pipeline: the next particle's loads are in flight while the
    // current one is processed
    size_t gi = c0 + threadIdx.x;
    bool cur_live = gi < c1;
    Loaded cur = load_particle(parts, nparts, P0, gi, cur_live);
#pragma unroll 1
    for (int it = 0; it < kChunk / kLdsThreads; it++) {
      const size_t gn = gi + kLdsThreads;
      const bool nxt_live = it + 1 < kChunk / kLdsThreads && gn < c1;
      const Loaded nxt = load_particle(parts, nparts, P0, gn, nxt_live);
      int nt = 0;
      int mock = mock0;
      if (cur_live) {
        nt = particle_tiles<true>(g, cur.r, cur.wscale, W, tiles, rec, flags);
        mock = cur.mock;
        if (nt) wmax = fmaxf(wmax, fabsf(rec.w));
      }
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const bool act = k < nt;
        if (!__ballot(act)) break;
        const bool here = act && mock == mock0;
        uint32_t slot = run_increment<true>(hist, here ? tiles[k] : 0u, here);
        if (act && !here) slot = atomicAdd(&cursors[(uint32_t)mock * tpm + tiles[k]], 1u);
        if (act) {
          if (slot < rec_cap) recs[slot] = rec;
          else atomicOr(flags, 1);
        }
      }
      cur = nxt;
      cur_live = nxt_live;
      gi = gn;
    }
  }
  publish_wmax(wmax, flags);
}

// ---- exclusive scan of uint32 counters (3 kernels) -------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanBlock = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v,
                                                         uint32_t *sh,
                                                         uint32_t &total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
      uint32_t t = sh[w];
      sh[w] = run;
      run += t;
    }
    sh[16] = run;
  }
  __syncthreads();
  total = sh[16];
  uint32_t res = x - v + sh[wid];
  __syncthreads();
  return res;
}

__global__ void __launch_bounds__(kScanThreads)
scan_local(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
           uint32_t *__restrict__ bsum, size_t n) {
  __shared__ uint32_t sh[17];
  size_t base = (size_t)blockIdx.x * kScanBlock + (size_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    v[i] = (base + i < n) ? in[base + i] : 0u;
    s += v[i];
  }
  uint32_t tot;
  uint32_t ex = block_exclusive_scan(s, sh, tot);
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    if (base + i < n) out[base + i] = ex;
    ex += v[i];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanThreads)
scan_sums(uint32_t *__restrict__ bsum, int nb) {
  __shared__ uint32_t sh[17];
  uint32_t carry = 0;
  for (int base = 0; base < nb; base += kScanThreads) {
    int i = base + threadIdx.x;
    uint32_t v = (i < nb) ? bsum[i] : 0u;
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(v, sh, tot);
    if (i < nb) bsum[i] = ex + carry;
    carry += tot;
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}

template <bool CURSORS>
__global__ void __launch_bounds__(kScanThreads)
scan_add(uint32_t *__restrict__ out, uint32_t *__restrict__ cursors,
         const uint32_t *__restrict__ bsum, size_t n, int nb) {
  size_t base = (size_t)blockIdx.x * kScanBlock;
  uint32_t off = bsum[blockIdx.x];
  for (int i = threadIdx.x; i < kScanBlock; i += kScanThreads) {
    size_t k = base + i;
    if (k < n) {
      uint32_t v = out[k] + off;
      out[k] = v;
      if (CURSORS) cursors[k] = v;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = bsum[nb];
}

__global__ void fill_zero_u32(uint32_t *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// ---- deposit: one workgroup per tile ------------------------------------------

__device__ __forceinline__ void window_weights(int scheme, float f, float *w) {
  switch (scheme) {
    case 0:
      w[0] = 1.f;
      break;
    case 1:
      w[0] = 1.f - f;
      w[1] = f;
      break;
    case 2: {
      float a = 0.5f - f;
      w[0] = 0.5f * a * a;
      w[1] = 0.75f - f * f;
      w[2] = 1.f - w[0] - w[1];
      break;
    }
    default: {
      float d2 = f * f;
      float w3 = d2 * f;
      float w2 = 1.f + 3.f * (f + d2 - w3);
      float w1 = 4.f - 6.f * d2 + 3.f * w3;
      w[3] = w3;
      w[2] = w2;
      w[1] = w1;
      w[0] = 6.f - w1 - w2 - w3;
      break;
    }
  }
}

// Fixed-point LDS accumulators (value * 2^s as a signed integer; FP32 LDS
// atomics are an order of magnitude slower on gfx950).  Integer addition is
// associative, so a tile does not depend on the order of its records.
//  - 32-bit (ds_add_u32): the exponent leaves about 2^-30 * cells-per-tile
//    relative precision per contribution -- ample for boxes, whose fields
//    are sums of positive weights;
//  - 64-bit (ds_add_u64): each contribution keeps its FP32 precision; used
//    when weights of both signs meet (survey data minus alpha * randoms),
//    where large-scale modes are small differences of large sums.
template <bool WIDE>
struct Acc;
// add(p, a, b, half) accumulates a * b rounded to an integer, half being
// 0.5 with the sign of the product.  32-bit: one FMA and a truncating
// conversion, trunc(a b + half) = a b rounded half away from zero (the
// rounding of a b + half in FP32 can move a product of 2^23 or more by one
// unit, a relative 1.2e-7, the FP32 precision the tile carries anyway).
template <>
struct Acc<false> {
  using T = int;
  static constexpr int TOP = 30;
  __device__ static __forceinline__ void add(int *p, float a, float b, float half) {
    atomicAdd(p, (int)__builtin_fmaf(a, b, half));
  }
};
template <>
struct Acc<true> {
  using T = long long;
  static constexpr int TOP = 62;
  __device__ static __forceinline__ void add(long long *p, float a, float b, float) {
    atomicAdd(reinterpret_cast<unsigned long long *>(p),
              (unsigned long long)__float2ll_rn(a * b));
  }
};

// Window axes x and y: an axis a >= branch_axes is deposited without
// branches (a window cell outside the tile gets weight 0 and the address of
// cell 0 along that axis; adding the integer 0 there is exact), an axis
// below keeps a branch per plane or row, which skips whole planes or rows
// outside the tile.  Measured on MI355X: y without branches for CIC and TSC,
// with them for PCS (whose 4-cell rows make a skipped row worth more).
// Axis z: the tile's rows are padded by W cells on each side, so a window
// row is W consecutive cells from one base address (constant LDS offsets,
// no per-cell mask or address arithmetic); cells in the padding are
// dropped, their tile deposits them from its own record.  A record's tile
// meets the union of the plain and the shifted window, so the plain
// window's first cell b lies in [T0 - W, T0 + T[2]) and the shifted one's
// in [T0 - W + 1, T0 + T[2]] (up to a wrap by Ng): W cells of padding on
// each side hold every window cell outside the tile.  When T[2] + W >= Ng
// (one tile, or two tiles on a tiny grid) a window offset is ambiguous
// modulo Ng, and z goes the branch-free way of x and y instead (ZPAD false).
template <int SCHEME>
constexpr int branch_axes() { return SCHEME == 3 ? 2 : 1; }

// Accumulate one particle's window into an LDS tile of fixed-point cells
// (rows of P2 = T[2] + 2 W cells, the tile's first z cell at W).
template <int SCHEME, bool WIDE, bool ZPAD>
__device__ __forceinline__ void deposit_one(typename Acc<WIDE>::T *lds, const int *T,
                                            const int *T0, int Ng, int P2, const int *b,
                                            const float *f, float w) {
  constexpr int W = SCHEME + 1;
  constexpr int kBranchAxes = branch_axes<SCHEME>();
  float wt[3][4];
  int l[3][4];
  unsigned m[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    window_weights(SCHEME, f[a], wt[a]);
    m[a] = 0;
    if (ZPAD && a == 2) break;
#pragma unroll
    for (int n = 0; n < W; n++) {
      int c = b[a] + n;
      if (c >= Ng) c -= Ng;
      const int lc = c - T0[a];
      const bool in = lc >= 0 && lc < T[a];
      if (a < kBranchAxes) {
        l[a][n] = lc;
        if (in) m[a] |= 1u << n;
      } else {
        l[a][n] = in ? lc : 0;
        if (!in) wt[a][n] = 0.f;
      }
    }
  }
  // ZPAD: offset in a padded row of the window's first z cell, b - T0 + W
  // modulo Ng, in [0, T[2] + W] (cells up to T[2] + 2W - 1 < P2).  The
  // difference itself lies in [W + T[2] - Ng, Ng + W): windows wrapped by Ng
  // at the tile at z = 0 (b in [Ng - W, Ng)) give b + W in [Ng, Ng + W),
  // and the shifted window of the last tile whose base wrapped to 0 gives
  // T[2] + W - Ng < 0; with T[2] + W < Ng one addition or subtraction of Ng
  // maps both into range and moves nothing else.
  int zoff = W;
  if (ZPAD) {
    zoff = b[2] - T0[2] + W;
    if (zoff >= Ng) zoff -= Ng;
    else if (zoff < 0) zoff += Ng;
  }
  if (SCHEME == 3) w *= (1.0f / 216.0f);
  const float half = __builtin_copysignf(0.5f, w);   // window weights are >= 0
  // row offsets: one integer multiply per window cell of x and of y
  int ry[4];
#pragma unroll
  for (int j = 0; j < W; j++) ry[j] = l[1][j] * P2 + zoff;
  const int PP = T[1] * P2;
#pragma unroll
  for (int i = 0; i < W; i++) {
    if (kBranchAxes > 0 && !(m[0] >> i & 1)) continue;
    const float wx = wt[0][i] * w;
    const int rx = l[0][i] * PP;
#pragma unroll
    for (int j = 0; j < W; j++) {
      if (kBranchAxes > 1 && !(m[1] >> j & 1)) continue;
      const float wxy = wx * wt[1][j];
      typename Acc<WIDE>::T *row = lds + rx + ry[j];
#pragma unroll
      for (int k = 0; k < W; k++)
        Acc<WIDE>::add(ZPAD ? row + k : row + l[2][k], wxy, wt[2][k], half);
    }
  }
}

// Fixed-point exponent of a tile: a cell value is a sum of at most cnt
// contributions of magnitude <= max|w| (window weights are <= 1 and sum to 1
// per axis), so |value| * 2^s < 2^TOP, one bit below the accumulator's sign.
template <int TOP>
__device__ __forceinline__ int tile_exponent(const int *flags, uint32_t cnt) {
  const float maxw = __uint_as_float((unsigned)flags[1]);
  const double bound = (double)maxw * (double)cnt * (1.0 + 1e-7);
  if (!(bound > 0.0)) return TOP;
  // ceil(log2(bound)) exactly: bound = m 2^e with m in [0.5, 1)
  int e;
  const double m = frexp(bound, &e);
  const int s = TOP - (m == 0.5 ? e - 1 : e);
  return s < -100 ? -100 : (s > 100 ? 100 : s);
}

typedef float v4f_t __attribute__((ext_vector_type(4)));

// Four consecutive mesh cells; non-temporal when the meshes are much larger
// than the caches (a template parameter: under a runtime flag the compiler
// merges both stores into an ordinary one).
template <bool NT>
__device__ __forceinline__ void store4(float *p, float4 v) {
  const v4f_t x = {v.x, v.y, v.z, v.w};
  if constexpr (NT) __builtin_nontemporal_store(x, reinterpret_cast<v4f_t *>(p));
  else *reinterpret_cast<v4f_t *>(p) = x;
}

// Deposit with the z transform fused in (boxes): a tile holds TX x TY whole
// z rows per grid, C lines over both grids, transformed by 1024 threads
// like the z pass; the rows (N + 2W cells) take about 130 KB of LDS at
// every size.  From N = 512, C = 32768 / N lines of 32-bit cells (16 values
// per thread).  Smaller grids put many more records into a tile, and a
// 32-bit cell's fixed-point step grows with the tile's record count (at
// N = 128 the quantisation reached 1e-5 of P(k) in sparsely populated
// bins): they take 64-bit cells, C = 16384 / N lines, 8 values per thread.
template <int N, bool IL>
struct ZfCfg {
  static constexpr bool WIDE = N <= 256;
  static constexpr int C = (WIDE ? 16384 : 32768) / N;   // lines per tile, both grids
  static constexpr int R = IL ? C / 2 : C;               // rows per grid
  static constexpr int TX = R >= 128 ? 16 : (R >= 32 ? 8 : 4);
  static constexpr int TY = R / TX;
  static constexpr int T = 1024;
  static_assert(TX * TY == R && TY >= 4, "tile rows");
};

template <int SCHEME, bool INTLACE, bool WIDE, bool NT, int NZ>
__global__ void __launch_bounds__(1024)
assign_deposit(PshAssignGeom g, const uint32_t *__restrict__ starts,
               const Rec *__restrict__ recs, size_t rec_cap,
               float *__restrict__ meshA, float *__restrict__ meshB,
               size_t mstride, const int *__restrict__ flags, size_t ntotal) {
  using AT = typename Acc<WIDE>::T;
  extern __shared__ __align__(16) long long ldsraw[];
  AT *ldsq = reinterpret_cast<AT *>(ldsraw);
  const int T[3] = {g.tile[0], g.tile[1], g.tile[2]};
  const int P2 = psh_tile_pitch(T[2], SCHEME);
  const int gcells = (int)psh_tile_cells(T, SCHEME);   // per grid, padded
  const bool zpad = T[2] + SCHEME + 1 < g.Ng;
  AT *ldsA = ldsq;
  AT *ldsB = ldsq + gcells;
  // tile index -> (mock, x, y, z) tile; tile counts need not be powers of
  // two (a grid of 3 * 2^k or 5 * 2^k cells has 3 or 5 times a power)
  const uint32_t tile = blockIdx.x;
  const uint32_t nz = (uint32_t)g.ntile[2], ny = (uint32_t)g.ntile[1];
  const uint32_t tpm = (uint32_t)g.ntile[0] * ny * nz;
  const int mock = (int)(tile / tpm);
  const uint32_t tl = tile - (uint32_t)mock * tpm;
  const uint32_t tyz = tl / nz;
  const int tzi = (int)(tl - tyz * nz);
  const int txi = (int)(tyz / ny);
  const int tyi = (int)(tyz - (uint32_t)txi * ny);
  const int T0[3] = {txi * T[0], tyi * T[1], tzi * T[2]};
  (void)ntotal;
  {
    // both grids are contiguous in LDS: clear them with 16-byte stores
    uint4 *z4 = reinterpret_cast<uint4 *>(ldsq);
    const int n4 = (int)(sizeof(AT) * gcells * (INTLACE ? 2 : 1) / 16);
    for (int i = threadIdx.x; i < n4; i += blockDim.x) z4[i] = make_uint4(0u, 0u, 0u, 0u);
  }
  __syncthreads();

  // records past the buffer were never written (overflow flag set, the run
  // is repeated with a larger buffer): never read beyond rec_cap
  const uint32_t cap = rec_cap < 0xffffffffu ? (uint32_t)rec_cap : 0xffffffffu;
  const uint32_t s0 = min(starts[tile], cap), s1 = min(starts[tile + 1], cap);
  const uint32_t cnt = s1 > s0 ? s1 - s0 : 0u;
  const int sexp = tile_exponent<Acc<WIDE>::TOP>(flags, cnt);
  const float scale = ldexpf(1.f, sexp);
  const float inv = ldexpf(1.f, -sexp);
  // Visit the bucket in the order i -> (i * S) mod cnt with S ~ 0.618 cnt and
  // gcd(S, cnt) = 1 (a permutation): neighbouring lanes take records far
  // apart in fill order, i.e. far apart in the (usually space-ordered)
  // catalogue, so their windows rarely share LDS banks.
  // (a bucket of at most one record per thread skips the search: the gcd
  // loop and 64-bit remainders would cost more than they save)
  uint32_t S = 1u;
  if (cnt > blockDim.x) {
    S = ((uint32_t)((float)cnt * 0.618034f)) | 1u;
    if (S >= cnt) S = 1u;
    for (;;) {
      uint32_t a = S, b = cnt;
      while (b) {
        uint32_t t = a % b;
        a = b;
        b = t;
      }
      if (a == 1u) break;
      S += 2u;
      if (S >= cnt) {
        S = 1u;
        break;
      }
    }
  }
  auto deposit_rec = [&](const Rec &r) {
    int b[3] = {(int)(r.bxy & 0xffffu), (int)(r.bxy >> 16), (int)(r.bz & 0xffffu)};
    const float ws = r.w * scale;
    auto dep = [&](AT *lds, const int *bb, const float *ff) {
      if (zpad) deposit_one<SCHEME, WIDE, true>(lds, T, T0, g.Ng, P2, bb, ff, ws);
      else deposit_one<SCHEME, WIDE, false>(lds, T, T0, g.Ng, P2, bb, ff, ws);
    };
    dep(ldsA, b, r.f);
    if (INTLACE) {
      const uint32_t dl = r.bz >> 16;
      int bs[3];
      float fs[3];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        int d = (dl >> a) & 1;
        bs[a] = b[a] + d;
        if (bs[a] >= g.Ng) bs[a] -= g.Ng;
        fs[a] = r.f[a] + 0.5f - (float)d;
      }
      dep(ldsB, bs, fs);
    }
  };
  if (cnt <= blockDim.x) {
    // at most one record per thread, on the fewest waves that hold them
    // (nw = ceil(cnt / 64)): the deposit is VALU-bound, so a wave issues as
    // many instructions for 20 live lanes as for 64, and the other waves go
    // straight to the barrier.  Lane l of wave w < nw takes record l * nw + w
    // (a bijection of [0, 64 nw)), so a wave's lanes hold records nw apart
    // in fill order rather than neighbours.
    const uint32_t nw = (cnt + 63u) >> 6, w = threadIdx.x >> 6;
    const uint32_t j = (threadIdx.x & 63u) * nw + w;
    if (w < nw && j < cnt) deposit_rec(recs[s0 + j]);
  } else {
    uint32_t x = (uint32_t)(((uint64_t)threadIdx.x * S) % cnt);
    const uint32_t D = (uint32_t)(((uint64_t)blockDim.x * S) % cnt);
    for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x) {
      const Rec r = recs[s0 + x];
      x += D;
      if (x >= cnt) x -= cnt;
      deposit_rec(r);
    }
  }
  __syncthreads();

  if constexpr (NZ > 0) {
    // The tile's rows, both grids, as NZ / 2-point complex lines (cells
    // 2n, 2n + 1 of a row), transformed like the z pass (z_r2c_kernel) and
    // split into the half-complex row X[0..NZ/2] (k_z <= kzlim stored):
    // X[k] = E + W^k O, X[M - k] = conj(E - W^k O).
    using namespace pshfft;
    using Z = ZfCfg<NZ, INTLACE>;
    constexpr int M = NZ / 2, C = Z::C, TT = Z::T, W = SCHEME + 1;
    using LZ = Lds<M, C, false>;
    constexpr int V = max_values<M, C, TT, false>();
    float2 v[V];
    pass_load<M, C, TT, 0, false>(v, [&](int c, int n) -> float2 {
      const AT *row = (c < Z::R ? ldsA + c * P2 : ldsB + (c - Z::R) * P2) + W;
      return make_float2((float)row[2 * n] * inv, (float)row[2 * n + 1] * inv);
    });
    __syncthreads();   // every cell read before the exchange tile overwrites them
    float *xl = reinterpret_cast<float *>(ldsq);
    Passes<M, C, TT, 0, false, false, 2>::run(v, xl, g.tw);
    constexpr int NP = M / 2 + 1;
    constexpr int NI = (C * NP + TT - 1) / TT;
    float rk[NI], rmk[NI];
    store_last<M, C, TT, false>(v, [&](int c, int n, float2 x) { xl[LZ::idx(c, n)] = x.x; });
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NI; it++) {
      const int item = threadIdx.x + it * TT;
      const int c = item / NP, k = item - c * NP;
      if (item < C * NP) {
        rk[it] = xl[LZ::idx(c, k)];
        rmk[it] = xl[LZ::idx(c, mirror<M>(k))];
      }
    }
    __syncthreads();
    store_last<M, C, TT, false>(v, [&](int c, int n, float2 x) { xl[LZ::idx(c, n)] = x.y; });
    __syncthreads();
    float2 *cA = reinterpret_cast<float2 *>(meshA + (size_t)mock * mstride);
    float2 *cB = INTLACE ? reinterpret_cast<float2 *>(meshB + (size_t)mock * mstride) : nullptr;
    const int H = g.row_pitch / 2;
#pragma unroll
    for (int it = 0; it < NI; it++) {
      const int item = threadIdx.x + it * TT;
      const int c = item / NP, k = item - c * NP;
      if (item >= C * NP) continue;
      const float ik = xl[LZ::idx(c, k)];
      const float imk = xl[LZ::idx(c, mirror<M>(k))];
      const float2 E = make_float2(0.5f * (rk[it] + rmk[it]), 0.5f * (ik - imk));
      const float2 O = make_float2(0.5f * (ik + imk), -0.5f * (rk[it] - rmk[it]));
      const float2 WO = cmul(g.tw[k], O);
      const int grid = c / Z::R, row = c - grid * Z::R;
      const int lx = row / Z::TY, ly = row - lx * Z::TY;
      float2 *orow = (grid ? cB : cA) + ((size_t)(T0[0] + lx) * g.Ng + (T0[1] + ly)) * H;
      if (k <= g.kzlim) mesh_store<NT>(orow + k, cadd(E, WO));
      if (M - k != k && M - k <= g.kzlim) {
        const float2 d = csub(E, WO);
        mesh_store<NT>(orow + (M - k), make_float2(d.x, -d.y));
      }
    }
    return;
  }

  // write the tile: rows of T[2] floats along z, 4 cells per thread
  float *outA = meshA + (size_t)mock * mstride;
  float *outB = INTLACE ? meshB + (size_t)mock * mstride : nullptr;
  const int nrow = T[0] * T[1];
  const int q4s = g.tile_shift[2] - 2;   // T[2] = 4 * 2^q4s (T[2] >= 4)
  const int q4 = 1 << q4s;
  for (int i = threadIdx.x; i < nrow * q4; i += blockDim.x) {
    const int row = i >> q4s, q = i & (q4 - 1);
    const int lx = row >> g.tile_shift[1], ly = row & (T[1] - 1);
    const size_t gidx = ((size_t)(T0[0] + lx) * g.Ng + (T0[1] + ly)) * g.row_pitch +
                        T0[2] + 4 * q;
    const int li = row * P2 + SCHEME + 1 + 4 * q;
    float4 va;
    va.x = (float)ldsA[li] * inv;
    va.y = (float)ldsA[li + 1] * inv;
    va.z = (float)ldsA[li + 2] * inv;
    va.w = (float)ldsA[li + 3] * inv;
    store4<NT>(outA + gidx, va);
    if (INTLACE) {
      float4 vb;
      vb.x = (float)ldsB[li] * inv;
      vb.y = (float)ldsB[li + 1] * inv;
      vb.z = (float)ldsB[li + 2] * inv;
      vb.w = (float)ldsB[li + 3] * inv;
      store4<NT>(outB + gidx, vb);
    }
  }
}

template <int SCHEME>
hipError_t launch_deposit(const PshAssignGeom &g, size_t ntiles,
                          const uint32_t *starts, const Rec *recs,
                          size_t rec_cap, float *meshA, float *meshB,
                          size_t mstride, const int *flags, size_t ntotal,
                          hipStream_t s) {
  const size_t lds = (g.wide ? sizeof(long long) : sizeof(int)) *
                     psh_tile_cells(g.tile, SCHEME) * (g.intlace ? 2 : 1);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  // non-temporal mesh stores when the launch's meshes outgrow the 256 MB
  // Infinity Cache (every cell is written once and read by the next pass
  // only after the whole launch)
  const bool stream = sizeof(float) * (size_t)g.tile[0] * g.tile[1] * g.tile[2] * ntiles *
                          (g.intlace ? 2 : 1) >
                      ((size_t)256 << 20);
  auto k = stream
               ? (g.wide ? (g.intlace ? assign_deposit<SCHEME, true, true, true, 0>
                                      : assign_deposit<SCHEME, false, true, true, 0>)
                         : (g.intlace ? assign_deposit<SCHEME, true, false, true, 0>
                                      : assign_deposit<SCHEME, false, false, true, 0>))
               : (g.wide ? (g.intlace ? assign_deposit<SCHEME, true, true, false, 0>
                                      : assign_deposit<SCHEME, false, true, false, 0>)
                         : (g.intlace ? assign_deposit<SCHEME, true, false, false, 0>
                                      : assign_deposit<SCHEME, false, false, false, 0>));
  if (g.zfuse) {
    // the fused z transform writes half-complex rows, 8 B per stored mode
    auto pick = [&](auto nz) {
      constexpr int NZ = decltype(nz)::value;
      constexpr bool W = ZfCfg<NZ, true>::WIDE;
      const int tx = g.intlace ? ZfCfg<NZ, true>::TX : ZfCfg<NZ, false>::TX;
      const int ty = g.intlace ? ZfCfg<NZ, true>::TY : ZfCfg<NZ, false>::TY;
      if (g.tile[0] != tx || g.tile[1] != ty || g.tile[2] != NZ || (g.wide != 0) != W) {
        k = nullptr;
        return;
      }
      k = stream ? (g.intlace ? assign_deposit<SCHEME, true, W, true, NZ>
                              : assign_deposit<SCHEME, false, W, true, NZ>)
                 : (g.intlace ? assign_deposit<SCHEME, true, W, false, NZ>
                              : assign_deposit<SCHEME, false, W, false, NZ>);
    };
    switch (g.Ng) {
      case 128: pick(std::integral_constant<int, 128>{}); break;
      case 256: pick(std::integral_constant<int, 256>{}); break;
      case 512: pick(std::integral_constant<int, 512>{}); break;
      case 1024: pick(std::integral_constant<int, 1024>{}); break;
      default: k = nullptr;
    }
    if (!k) return hipErrorInvalidValue;
  }
  if (lds > 64 * 1024)
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
  // one thread per 16 cells, 256..1024: a large tile fills its CU alone (its
  // LDS admits no second workgroup), so it brings its own waves
  const size_t tcells = (size_t)g.tile[0] * g.tile[1] * g.tile[2];
  unsigned threads = 256;
  while (threads < 1024 && threads * 16 < tcells) threads *= 2;
  if (g.zfuse) threads = 1024;   // the fused transform's block shape
  hipLaunchKernelGGL(k, dim3((unsigned)ntiles), dim3(threads), lds, s, g, starts, recs, rec_cap,
                     meshA, meshB, mstride, flags, ntotal);
  return hipGetLastError();
}

}  // namespace

size_t psh_assign_record_bytes() { return sizeof(Rec); }

bool psh_zfuse_wide(int Ng) { return Ng <= 256; }

bool psh_zfuse_tile(int Ng, int intlace, int tile[3]) {
  auto set = [&](auto nz) {
    constexpr int NZ = decltype(nz)::value;
    tile[0] = intlace ? ZfCfg<NZ, true>::TX : ZfCfg<NZ, false>::TX;
    tile[1] = intlace ? ZfCfg<NZ, true>::TY : ZfCfg<NZ, false>::TY;
    tile[2] = NZ;
    return true;
  };
  switch (Ng) {
    case 128: return set(std::integral_constant<int, 128>{});
    case 256: return set(std::integral_constant<int, 256>{});
    case 512: return set(std::integral_constant<int, 512>{});
    case 1024: return set(std::integral_constant<int, 1024>{});
    default: return false;
  }
}

size_t psh_assign_whist_words(const PshAssignGeom &g, size_t ntotal) {
  const size_t tpm = (size_t)g.ntile[0] * g.ntile[1] * g.ntile[2];
  if (tpm > (size_t)kLdsTiles || ntotal == 0) return 0;
  const size_t blocks = (ntotal + kChunk - 1) / kChunk;
  const size_t words = blocks * tpm;
  return words <= ((size_t)1 << 30) ? words : 0;   // at most 4 GB of histograms
}

hipError_t psh_fill_zero_u32(uint32_t *p, size_t n, hipStream_t s) {
  unsigned blocks = (unsigned)((n + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(fill_zero_u32, dim3(blocks), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t psh_exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint32_t *tmp, size_t n,
                                  hipStream_t s) {
  if (n == 0) return hipErrorInvalidValue;
  const int nsb = (int)((n + kScanBlock - 1) / kScanBlock);
  hipLaunchKernelGGL(scan_local, dim3(nsb), dim3(kScanThreads), 0, s, in, out, tmp, n);
  hipLaunchKernelGGL(scan_sums, dim3(1), dim3(kScanThreads), 0, s, tmp, nsb);
  hipLaunchKernelGGL(scan_add<false>, dim3(nsb), dim3(kScanThreads), 0, s, out,
                     (uint32_t *)nullptr, tmp, n, nsb);
  return hipGetLastError();
}

size_t psh_assign_scan_bytes(size_t ntiles) {
  size_t nsb = (ntiles + kScanBlock - 1) / kScanBlock;
  return (nsb + 1) * sizeof(uint32_t);
}

hipError_t psh_assign_run(const PshAssignGeom &g, const PshDevPart *d_parts,
                          int nparts, size_t ntotal, int nmock,
                          uint32_t *counts, uint32_t *starts, uint32_t *cursors,
                          void *records, size_t rec_cap, float *meshA,
                          float *meshB, size_t mstride, uint32_t *d_scan_tmp,
                          int *d_flags, hipStream_t s, hipEvent_t ev_deposit,
                          uint32_t *whist) {
  if (nparts < 1) return hipErrorInvalidValue;
  const size_t ntiles = (size_t)g.ntile[0] * g.ntile[1] * g.ntile[2] * nmock;
  const int nsb = (int)((ntiles + kScanBlock - 1) / kScanBlock);
  hipError_t e;
  if ((e = psh_fill_zero_u32(counts, ntiles, s)) != hipSuccess) return e;
  unsigned pblocks = (unsigned)((ntotal + 255) / 256);
  if (pblocks > 16384) pblocks = 16384;
  if (pblocks == 0) pblocks = 1;
  Rec *recs = reinterpret_cast<Rec *>(records);
  const size_t tpm = (size_t)g.ntile[0] * g.ntile[1] * g.ntile[2];
  const bool lds = tpm <= (size_t)kLdsTiles;
  const unsigned cblocks = (unsigned)((ntotal + kChunk - 1) / kChunk);
  const size_t lds_bytes = sizeof(uint32_t) * tpm;
  if (lds && lds_bytes > 64 * 1024) {
    (void)hipFuncSetAttribute((const void *)assign_bucket_lds<false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    (void)hipFuncSetAttribute((const void *)assign_bucket_lds<true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  }
  if (ntotal > 0) {
    if (lds)
      hipLaunchKernelGGL((assign_bucket_lds<false>), dim3(cblocks), dim3(kLdsThreads),
                         lds_bytes, s, g, d_parts, nparts, ntotal, counts, cursors, recs,
                         rec_cap, d_flags, whist);
    else
      hipLaunchKernelGGL((assign_bucket<false>), dim3(pblocks), dim3(256), 0, s, g,
                         d_parts, nparts, ntotal, counts, cursors, recs, rec_cap,
                         d_flags);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(scan_local, dim3(nsb), dim3(kScanThreads), 0, s, counts,
                     starts, d_scan_tmp, ntiles);
  hipLaunchKernelGGL(scan_sums, dim3(1), dim3(kScanThreads), 0, s, d_scan_tmp, nsb);
  hipLaunchKernelGGL(scan_add<true>, dim3(nsb), dim3(kScanThreads), 0, s, starts,
                     cursors, d_scan_tmp, ntiles, nsb);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (ntotal > 0) {
    if (lds)
      hipLaunchKernelGGL((assign_bucket_lds<true>), dim3(cblocks), dim3(kLdsThreads),
                         lds_bytes, s, g, d_parts, nparts, ntotal, counts, cursors, recs,
                         rec_cap, d_flags, whist);
    else
      hipLaunchKernelGGL((assign_bucket<true>), dim3(pblocks), dim3(256), 0, s, g,
                         d_parts, nparts, ntotal, counts, cursors, recs, rec_cap,
                         d_flags);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (ev_deposit) (void)hipEventRecord(ev_deposit, s);
  switch (g.scheme) {
    case 0: return launch_deposit<0>(g, ntiles, starts, recs, rec_cap, meshA, meshB, mstride, d_flags, ntotal, s);
    case 1: return launch_deposit<1>(g, ntiles, starts, recs, rec_cap, meshA, meshB, mstride, d_flags, ntotal, s);
    case 2: return launch_deposit<2>(g, ntiles, starts, recs, rec_cap, meshA, meshB, mstride, d_flags, ntotal, s);
    default: return launch_deposit<3>(g, ntiles, starts, recs, rec_cap, meshA, meshB, mstride, d_flags, ntotal, s);
  }
}
