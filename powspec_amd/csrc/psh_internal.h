// psh_internal.h -- shared declarations of the gfx950 kernels and the host
// orchestration of the mesh -> FFT -> P_l(k) path.
//
// Memory layout of one mesh in HBM (Ng = grid size, M = Ng/2):
//   real field     [Ng][Ng][2H] float,  rows padded to 2H floats
//   complex field  [Ng][Ng][H]  float2, same bytes (in-place r2c),
//                  kz = 0..M stored (M+1 entries), H = M + PSH_PAD
// The padding keeps every 16-column kz tile 128-byte aligned and lets the
// z-direction r2c run in place.  x is the slowest index and z the fastest,
// as in the reference (src/define.h:134).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#define PSH_PAD 16          // complex padding per row (one 16-column tile)
#define PSH_MAX_NL 7
// fields of one survey pole l > 0 that one x pass bins together (2 l + 1 <= 13)
#define PSH_MAX_YLM_FIELDS 13

// Per-axis FP64 tables used by the binning epilogues, device resident.
// For axis a and index i (signed frequency i' for x,y; i for z):
//   kv[a][i] = vec_a * i'          (vec_a = 2 pi / L_a)
//   ksq[a][i] = kv * kv
//   al[a][i] = alias correction factor (src/multipole.c:46-100)
//   kd[a][i] = i' / L_a            (direction of k for Y_lm(k))
struct PshAxisTables {
  const double *kv[3];
  const double *ksq[3];
  const double *al[3];
  const double *kd[3];
  // FP32 roundings of kv, ksq, al for the per-mode fast path
  const float *kvf[3];
  const float *ksqf[3];
  const float *alf[3];
  float almax[3];   // largest alias factor per axis
};

// Binning geometry (src/multipole.c:306-421).
struct PshBins {
  int nbin;
  int logscale;
  double kedge0;   // kedge[0]
  double kedgeN;   // kedge[nbin]
  double dk;
  int nl;
  int poles[PSH_MAX_NL];
  double los[3];
  // FP32 fast path: t = (|k| - kedge0) / dk decided in FP32 when it lies
  // farther than an error bound from an integer, else recomputed in FP64
  float kedge0f, invdkf, k0dk;
  float losf[3];
  // boxes: |mu| cells per bin of the fused x pass's sums (1: none); cell c
  // holds the modes with |mu| in [c / nmu, (c + 1) / nmu)
  int nmu;
};

// What the fused last FFT pass does with each mode.
enum PshEpilogue {
  PSH_EPI_SIM = 0,        // sim: |F|^2 alias mult L_l(mu), DC skipped
  PSH_EPI_SURVEY0 = 1,    // survey l=0: |F|^2 alias mult, DC kept; F stored
  PSH_EPI_SURVEYL = 2,    // survey l>0: Re(F0 conj(Y_lm(k) S)) alias mult
  PSH_EPI_STORE = 3,      // plain transform, result stored
};

// ---------------------------------------------------------------------------
// Assignment (k_assign.hip)
// ---------------------------------------------------------------------------
struct PshAssignGeom {
  int Ng;
  int scheme;           // 0..3
  int intlace;          // second (half-cell shifted) grid
  int wide;             // 64-bit fixed-point tiles (weights of both signs)
  int check_box;        // count pass flags (bit 4) coordinates outside [bmin, bmin + L)
  int tile[3];          // tile extent per axis (a power of two)
  int tile_shift[3];    // log2(tile)
  int ntile[3];         // Ng / tile
  int row_pitch;        // floats per mesh row (2H)
  double bmin[3];       // box corner of the plain grid
  double smin[3];       // box corner of the shifted grid
  double L[3];          // box size
  // z transform fused into the deposit (tiles of whole z rows, boxes): the
  // deposit writes half-complex rows (k_z <= kzlim) instead of real ones
  int zfuse;
  int kzlim;
  const float2 *tw;     // exp(-2 pi i m / Ng), m < Ng
};

// Tile of the deposit with the fused z transform for a grid size: whole z
// rows, tx x ty rows per grid; false when this build has no fused deposit
// for Ng.  psh_zfuse_wide: whether it accumulates 64-bit cells.
bool psh_zfuse_tile(int Ng, int intlace, int tile[3]);
bool psh_zfuse_wide(int Ng);

// Deposit tile in LDS, per grid: rows along z padded by the window width
// W = scheme + 1 on each side (window cells past the tile's z range land
// there and are dropped; the neighbouring tile deposits them from its own
// record), the grid's cell count rounded up to 4 (16-byte clears).
__host__ __device__ inline int psh_tile_pitch(int tz, int scheme) { return tz + 2 * (scheme + 1); }
__host__ __device__ inline size_t psh_tile_cells(const int t[3], int scheme) {
  return ((size_t)t[0] * t[1] * psh_tile_pitch(t[2], scheme) + 3) & ~(size_t)3;
}

// One catalogue part of a (possibly batched) assignment; the table lives in
// device memory (uploaded before any graph capture).
struct PshDevPart {
  const double *xyzw;   // device psh_data records (x,y,z,w)
  size_t n;             // number of records
  size_t start;         // first global particle index of this part
  float wscale;         // 1 for data, -alpha for randoms
  int mock;             // destination mesh pair index
};

size_t psh_assign_record_bytes();
// Bytes of scan scratch for ntiles counters.
size_t psh_assign_scan_bytes(size_t ntiles);

// Count the (particle, tile) pairs, fill tile buckets and deposit the plain
// (meshA) and shifted (meshB, interlacing only) meshes of nmock mocks; mock
// m's meshes start at meshA + m*mstride, meshB + m*mstride.
// counts, starts (ntiles+1), cursors: uint32 scratch; records: rec_cap
// records; flags: bit 0 record overflow, bit 1 inconsistent shifted base.
hipError_t psh_assign_run(const PshAssignGeom &g, const PshDevPart *d_parts,
    int nparts, size_t ntotal, int nmock, uint32_t *counts, uint32_t *starts,
    uint32_t *cursors, void *records, size_t rec_cap, float *meshA,
    float *meshB, size_t mstride, uint32_t *d_scan_tmp, int *d_flags,
    hipStream_t s, hipEvent_t ev_deposit = nullptr, uint32_t *whist = nullptr);

// uint32 words of the per-workgroup tile histograms that let the fill pass
// skip its recount (0 when the LDS bucketing does not apply or they would
// exceed 4 GB); whist passed to psh_assign_run must hold that many.
size_t psh_assign_whist_words(const PshAssignGeom &g, size_t ntotal);

// ---------------------------------------------------------------------------
// FFT (k_fft.hip)
// ---------------------------------------------------------------------------

// Twiddle table exp(-2 pi i m / N), m = 0..N-1, float2, from FP64:
// allocated (psh_fft_twiddles) or written into d_tw's N entries.
hipError_t psh_fft_twiddles(int N, float2 **d_tw);
hipError_t psh_fft_twiddles_fill(int N, float2 *d_tw);

// Forward r2c along z, in place or out of place, for nmesh meshes laid out
// contiguously (mesh stride mstride floats).  ylm: optional Y_lm(r) weight
// (ell > 0) of the survey estimator with the reference's first-row rule.
struct PshYlmR {
  int ell, m;            // ell = 0: no weighting
  float norm;            // normalisation of the real-form Y_lm
  // r-space position of cell i on axis a: org[a] + i step[a], the cell's
  // (i + smin_g) L (a common factor of all three axes leaves Y_lm unchanged)
  float org[3], step[3];
};
// tw: twiddle table of N = Ng entries.
// kzlim: outputs past k_z = kzlim are not stored (fields that only the
// fused x pass reads, whose tiles past it hold no binned mode).
hipError_t psh_fft_z_r2c(int Ng, const float *in, float2 *out, int nmesh,
    size_t mstride_floats, const float2 *tw, const PshYlmR &ylm,
    hipStream_t s, int kzlim = 1 << 30);

// Inverse c2r along z in place, output scaled by `scale`.
hipError_t psh_fft_z_c2r(int Ng, float2 *inout, int nmesh,
    size_t mstride_floats, const float2 *tw, float scale, hipStream_t s);

// Complex FFT along y (axis 1) in place; dir +1 inverse, -1 forward.
// ylive (forward y pass of fields that only the fused x pass reads, or
// nullptr): per x-pass tile of k_z columns, the largest |k_y| index whose
// tile holds binned modes; outputs past it are not stored.
// sched: schedule of the column pass (0 default; see psh_fft_y_sched_ok).
hipError_t psh_fft_y(int Ng, float2 *data, int nmesh, size_t mstride_floats,
    const float2 *tw, int inverse, hipStream_t s, const int *ylive = nullptr,
    int sched = 0);
// Schedules of the column pass: 0..PSH_FFT_Y_SCHEDS-1; whether one is valid
// at a grid size (its LDS fits and it differs from schedule 0).
#define PSH_FFT_Y_SCHEDS 6
int psh_fft_y_sched_ok(int Ng, int sched);

// Complex FFT along x (axis 0) from `in` to `out` (may alias); no epilogue.
hipError_t psh_fft_x(int Ng, const float2 *in, float2 *out, int nmesh,
    size_t mstride_floats, const float2 *tw, int inverse, hipStream_t s);

// Forward FFT along x fused with the binning epilogue.
struct PshXBinArgs {
  int epi;                 // PshEpilogue
  const float2 *A;         // plain-grid field (per mock stride mstride)
  const float2 *B;         // shifted-grid field or nullptr
  float2 *store;           // F store target for SURVEY0 (usually == A)
  const float2 *F0;        // SURVEYL: the l=0 field
  // SURVEYL: nf fields of Y_lm-weighted densities (A + f fstride_floats),
  // field f binned with the weight Y_{ells[f] ms[f]}(k), normalisation
  // ynorms[f]; their sum is binned once
  int nf;
  size_t fstride_floats;
  int ells[PSH_MAX_YLM_FIELDS], ms[PSH_MAX_YLM_FIELDS];
  float ynorms[PSH_MAX_YLM_FIELDS];
  float vec0f;             // 2 pi / L_x (FP32)
  int nmock;
  size_t mstride_floats;
  PshAxisTables ax;
  PshBins bins;
  const float2 *phase;     // interlace phase tables [3][Ng] exp(i pi i'/Ng)
  double *partials;        // [grid][nacc] per workgroup partial sums
  int grid;                // number of persistent workgroups per mock
  const int *ylive;        // per k_z tile, the largest |k_y| with binned modes
                           // (tiles past it are skipped), or nullptr: all tiles
  int kb0, nbw;            // bins [kb0, kb0 + nbw) of this launch (nbw 0: all);
                           // partials rows hold every bin, a launch writes its own
  int has_dc;              // the k = 0 mode of the field is dc
  double dc;               // (sum of its particles' weights, exact in FP64)
};
int psh_xbin_nacc(const PshBins &b, int epi);
// k_z columns per tile of the fused x pass at a grid size.
int psh_xbin_tile_cols(int Ng);
// LDS bytes of the fused x pass for a grid size, interlacing and number of
// accumulators (0 for an unsupported grid size); at most 160 KB run.
// Grid sizes with compiled transforms: powers of two 4..2048, 3 * 2^k
// (48..1536) and 5 * 2^k (80..1280).
bool psh_grid_supported(int Ng);
size_t psh_xbin_lds_bytes(int Ng, bool il, int nacc);
// Most bins one launch of the fused x pass holds in LDS with per_bin
// accumulators per bin (more bins take several launches over bin windows).
int psh_xbin_window_bins(int Ng, bool il, int per_bin);
// Workgroups of the fused x pass one CU holds at once for a window of nacc
// accumulators (0 when the occupancy query fails).
int psh_xbin_per_cu(int Ng, int epi, bool il, int nacc);
hipError_t psh_fft_x_bin(int Ng, const PshXBinArgs &a, const float2 *tw,
    hipStream_t s);

// ---------------------------------------------------------------------------
// Geometry and reductions (k_bin.hip)
// ---------------------------------------------------------------------------

// cnt (uint64), km (sum of kmod*mult) and lcnt (sum of mult*L_l(mu), DC
// skipped) per bin into out_cnt, out_km, out_lcnt[nl*nbin]; the sums are
// exact integer or fixed-point sums per workgroup added in a fixed order
// (the same on every run).  scratch: psh_geometry_scratch_bytes(b) bytes.
size_t psh_geometry_scratch_bytes(const PshBins &b);
hipError_t psh_geometry(int Ng, const PshAxisTables &ax, const PshBins &b,
    unsigned long long *out_cnt, double *out_km, double *out_lcnt, void *scratch,
    hipStream_t s);

// Sum partials[nmock][grid][nacc] over grid in a fixed order into
// out[nmock][nacc].
hipError_t psh_reduce_partials(const double *partials, int nmock, int grid,
    int nacc, double *out, hipStream_t s);

// Exclusive scan of n uint32 counts into out[0..n] (out[n] = total); tmp
// holds psh_assign_scan_bytes(n) bytes.  in and out must not overlap.
hipError_t psh_exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint32_t *tmp,
                                  size_t n, hipStream_t s);

// Fill a float buffer with zeros (graph-safe replacement of memset).
hipError_t psh_fill_zero_u32(uint32_t *p, size_t n, hipStream_t s);

// Min/max of the coordinates of n psh_data records (FP64, exact) and the
// sum of their weights (FP64, a fixed order per block).
// out: [blocks][PSH_BOUNDS_STRIDE] partial (min x,y,z, max x,y,z, sum w).
#define PSH_BOUNDS_STRIDE 7
hipError_t psh_coord_bounds(const double *xyzw, size_t n, double *partial,
    int blocks, hipStream_t s);

// ---------------------------------------------------------------------------
// Learned alias correction (k_mlp.hip)
// ---------------------------------------------------------------------------

// Feature rows X[nbin * nmu][32] of a plain box run from its raw monopole
// cell sums cells[nbin][nmu] (device), mode counts cnt[nbin] (device), the
// normalisation, shot noise, k_Ny and binning of the run, the scheme id.
hipError_t psh_alias_features(const double *cells, const unsigned long long *cnt, int nbin,
    int nmu, double norm, double shot, double kny, double kedge0, double dk, int logscale,
    int scheme, float *X, double *scratch, hipStream_t s);
// Doubles of device scratch psh_alias_features needs for nbin bins.
size_t psh_alias_scratch_doubles(int nbin);
// out[l][b] = sum_c exp(mean_n Y[n][b nmu + c]) cells[l][b][c] (device).
hipError_t psh_alias_apply(const float *Y, int nnet, const float *X, const double *cells, int nl,
    int nbin, int nmu, double *out, hipStream_t s);
// Network outputs Y[n][B] of rows X[B][32] for the nnet networks whose
// parameters follow one another from d_params (psh_mlp_param_count() each),
// one launch on stream s (no synchronisation).
hipError_t psh_mlp_forward_async(const float *d_params, const float *d_X, int B, float *d_Y,
    hipStream_t s, int nnet = 1);
