/*
 * powspec_hip.h -- C-ABI of the MI355X (gfx950) mesh -> FFT -> P_l(k) library.
 *
 * Two layers:
 *
 *  1. Drop-in layer: psh_genr_mesh / psh_powspec / psh_mesh_destroy /
 *     psh_powspec_destroy take the place of the reference's genr_mesh and
 *     powspec (src/genr_mesh.h:86, src/multipole.h:78, called from
 *     src/powspec.c:47,55).  psh_data and psh_cata have the exact memory
 *     layout of the reference's DATA and CATA (src/read_cata.h:42-59), so a
 *     CATA* can be passed as a psh_cata*.  psh_conf carries the CONF fields the
 *     path reads (src/load_conf.h:40-93).  NULL is returned on error, with a
 *     message on stderr, like the reference.
 *
 *  2. Context layer: psh_ctx_* functions with integer status codes, device-
 *     resident particle input, batched mocks and stage timings.  The drop-in
 *     layer is built on it.
 *
 * All pointers are plain host or device pointers; no framework types.
 * Status codes follow the reference's families (src/define.h:112-124).
 */
#ifndef POWSPEC_HIP_H
#define POWSPEC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define PSH_OK 0
#define PSH_ERR_MEMORY (-1)      /* POWSPEC_ERR_MEMORY (src/define.h:112)  */
#define PSH_ERR_ARG (-2)         /* POWSPEC_ERR_ARG    (src/define.h:113)  */
#define PSH_ERR_FILE (-3)        /* POWSPEC_ERR_FILE   (src/define.h:114)  */
#define PSH_ERR_CATA (-11)       /* POWSPEC_ERR_CATA   (src/define.h:119)  */
#define PSH_ERR_CNVT (-12)       /* POWSPEC_ERR_CNVT   (src/define.h:120)  */
#define PSH_ERR_MESH (-13)       /* POWSPEC_ERR_MESH   (src/define.h:121)  */
#define PSH_ERR_PK (-14)         /* POWSPEC_ERR_PK     (src/define.h:122)  */
#define PSH_ERR_DEVICE (-20)     /* HIP runtime failure (no reference code) */
#define PSH_ERR_UNSUPPORTED (-21)/* e.g. GRID_SIZE not a power of two       */

#define PSH_ASSIGN_NGP 0         /* powspec_assign_t (src/genr_mesh.h:40-45) */
#define PSH_ASSIGN_CIC 1
#define PSH_ASSIGN_TSC 2
#define PSH_ASSIGN_PCS 3

#define PSH_MAX_ELL 6            /* POWSPEC_MAX_ELL (src/define.h:96) */

/* ---- data types ------------------------------------------------------------ */

/* Same layout as the reference's DATA (src/read_cata.h:42-45). */
typedef struct {
  double x[3];
  double w;
} psh_data;

/* Same layout as the reference's CATA (src/read_cata.h:48-59). */
typedef struct {
  int num;
  psh_data **data;
  psh_data **rand;
  size_t *ndata;
  size_t *nrand;
  double *wdata;
  double *wrand;
  double *alpha;
  double *shot;
  double *norm;
} psh_cata;

/* The CONF fields the path reads (src/load_conf.h:40-93), with kmin/kmax as
 * the reference holds them after load_conf: log10-converted when logscale is
 * set (src/load_conf.c:1351-1355); kmax <= 0 means unset.  poles must be
 * sorted and unique (src/load_conf.c:1056-1074). */
typedef struct {
  int ndata;              /* number of data catalogues: 1 or 2          */
  int issim;              /* CUBIC_SIM                                   */
  int gsize;              /* GRID_SIZE                                   */
  int assign;             /* PARTICLE_ASSIGN                             */
  int intlace;            /* GRID_INTERLACE                              */
  int npole;              /* number of entries of poles                  */
  int poles[PSH_MAX_ELL + 1];
  int has_bsize;          /* BOX_SIZE given (reference: bsize != NULL)   */
  double bsize[3];        /* BOX_SIZE                                    */
  double bpad[3];         /* BOX_PAD                                     */
  double los[3];          /* LINE_OF_SIGHT                               */
  int logscale;           /* LOG_SCALE                                   */
  double kmin, kmax, kbin;/* KMIN, KMAX, BIN_SIZE                        */
  int isauto[2];
  int iscross;
  int verbose;            /* VERBOSE                                     */
} psh_conf;

/* Result of one catalogue set, equivalent to the reference's PK
 * (src/multipole.h:38-61).  Arrays are owned by the library (drop-in layer)
 * or by the caller (context layer, sized with psh_ctx_nbin). */
typedef struct {
  int nbin;
  int nl;
  int poles[PSH_MAX_ELL + 1];
  double *k;          /* [nbin]      bin centres                      */
  double *kedge;      /* [nbin + 1]  bin edges                        */
  double *km;         /* [nbin]      mean |k| of the modes in the bin */
  uint64_t *cnt;      /* [nbin]      number of modes                  */
  double *lcnt;       /* [nl*nbin]   sum of mult*L_l(mu) (boxes)      */
  double *pl[2];      /* [nl*nbin]   auto multipoles, l-major         */
  double *xpl;        /* [nl*nbin]   cross multipoles                 */
} psh_pk;

/* Mesh metadata kept after psh_powspec, as the reference's MESH keeps
 * Ng, box and scheme for the output writer (src/save_res.c:70-84). */
typedef struct psh_mesh psh_mesh;

/* ---- drop-in layer ------------------------------------------------------- */

/* Replaces genr_mesh (src/genr_mesh.c:874-926): defines the box, assigns the
 * catalogues to device meshes, sets shot/norm for simulation boxes.  Like the
 * reference it frees cat->data[i] and cat->rand[i] (which must come from
 * malloc) and sets cat->data and cat->rand to NULL.  NULL on error. */
psh_mesh *psh_genr_mesh(const psh_conf *conf, psh_cata *cat);

/* Replaces powspec (src/multipole.c:1179-1278): FFTs, interlacing, binning,
 * normalisation.  Releases the device meshes; the psh_mesh keeps its
 * metadata.  NULL on error. */
psh_pk *psh_powspec(const psh_conf *conf, const psh_cata *cat, psh_mesh *mesh);

/* Replace mesh_destroy (src/genr_mesh.c:615-640) and powspec_destroy
 * (src/multipole.c:270-295). */
void psh_mesh_destroy(psh_mesh *mesh);
void psh_powspec_destroy(psh_pk *pk);

/* Mesh metadata accessors (for the output writer, src/save_res.c:70-84). */
int psh_mesh_info(const psh_mesh *mesh, int *Ng, double bsize[3],
    double bmin[3], int *assign, int *intlace);

/* ---- output files ---------------------------------------------------------- */

/* Mesh metadata the output header reports (src/save_res.c:70-84). */
typedef struct {
  int Ng;
  double bsize[3];
  double bmin[3];
  int assign;
  int intlace;
} psh_mesh_meta;

/* Writes the reference's output files (src/save_res.c:35-230): one auto
 * power spectrum file per catalogue with isauto set (fauto[c]) and, when
 * iscross is set, the cross spectrum file fcross.  header != 0 adds the
 * catalogue, mesh and normalisation lines (OUTPUT_HEADER).  Numbers are
 * printed with the reference's formats (%.10lg, %lg for the box, %zu for
 * mode counts), so the files are byte-identical for identical values.
 * Returns PSH_OK, PSH_ERR_ARG or PSH_ERR_FILE. */
int psh_write_pk(const psh_conf *conf, const psh_cata *cat, const psh_mesh_meta *meta,
    const psh_pk *pk, const char *const fauto[2], const char *fcross, int header);

/* Replaces save_res (src/save_res.c:35): psh_write_pk with the metadata of a
 * psh_mesh kept by psh_powspec. */
int psh_save_res(const psh_conf *conf, const psh_cata *cat, const psh_mesh *mesh,
    const psh_pk *pk, const char *const fauto[2], const char *fcross, int header);

/* ---- survey coordinate conversion ------------------------------------------ */

/* The CONF fields of the conversion (src/load_conf.h:67-72): cosmology with
 * omega_k = 1 - omega_m - omega_l as load_conf sets it, DE_EOS_W,
 * CMVDST_ERR, and the Z_CMVDST_CONV samples (already read, host arrays;
 * NULL/0 selects the Gauss-Legendre integration), plus DATA_CONVERT /
 * RAND_CONVERT per catalogue. */
typedef struct {
  double omega_m, omega_l, omega_k;
  double eos_w;
  double ecdst;
  const double *zsp, *dsp;
  size_t nsp;
  int dcnvt[2], rcnvt[2];
} psh_cnvt;

/* Replaces cnvt_coord (src/cnvt_coord.c:551-606): (RA [deg], Dec [deg],
 * redshift) of the flagged catalogues -> comoving (x, y, z) in place, on the
 * GPU, in FP64.  Integration: the Gauss-Legendre order is chosen by the
 * reference's convergence test over the catalogues' redshift range
 * (src/cnvt_coord.c:138-239, 306-341, 494-512); negative redshifts are an
 * error.  Interpolation: natural cubic spline of the samples with the
 * reference's binary search; redshifts outside [z_0, z_last) give HUGE_VAL
 * distances as in the reference.  Returns PSH_OK or PSH_ERR_CNVT /
 * PSH_ERR_ARG / PSH_ERR_DEVICE. */
int psh_cnvt_coord(const psh_cnvt *cv, psh_cata *cat);

/* The same on device-resident records: narr arrays d_data[i] of n[i]
 * records, all converted with one order.  *order_out (may be NULL) receives
 * the integration order (0 for interpolation). */
int psh_cnvt_device(const psh_cnvt *cv, psh_data *const *d_data, const size_t *n,
    int narr, int *order_out);

/* ---- synthetic catalogues ------------------------------------------------- */
/* No counterpart in the reference: seeded generators for the benchmark and
 * test workloads.  All draw from Philox4x32-10 streams (counter = (index,
 * stream, draw), key = seed) that powspec_amd/synth.py reproduces on the
 * host, so CPU and GPU generators see the same random numbers. */

/* The four Philox words of counters (i, stream, draw), i < n -> d_out[4i..]. */
int psh_gen_philox(uint32_t *d_out, size_t n, uint64_t seed, uint32_t stream, uint32_t draw);

/* n records uniform in [0, L)^3 (kept below the box edge), unit weights. */
int psh_gen_uniform_box(psh_data *d_out, size_t n, const double L[3], uint64_t seed);

/* Clustered periodic box: Gaussian field with P(k) = A (k/0.1)^nidx
 * exp(-(k/kc)^2) on an ng^3 grid, lognormal, Poisson sampled at nbar (or at
 * npart expected particles when npart > 0), uniform in the cell, Zel'dovich
 * shift f Psi_z along z, periodic.  *d_out is allocated by the library (free
 * with psh_device_free); *n_out receives the particle count. */
typedef struct {
  int ng;
  double L;
  double nbar, npart;
  double f;
  double A, nidx, kc;
  uint64_t seed;
} psh_box_spec;
int psh_gen_clustered_box(const psh_box_spec *g, psh_data **d_out, size_t *n_out);

/* Survey sky sample: RA uniform in ra[], sin(Dec) uniform in dec[] (degrees),
 * redshift from a Gaussian (zmu, zsigma) truncated to [zlo, zhi); records
 * (RA, Dec, z, w_fkp) with n(z) = nbar0 exp(-(z-zmu)^2 / (2 zsigma^2)) and
 * w_fkp = 1 / (1 + n(z) p0); d_nz (may be NULL) receives n(z). */
typedef struct {
  double ra[2], dec[2];
  double zmu, zsigma, zlo, zhi;
  double nbar0, p0;
  uint64_t seed;
} psh_sky_spec;
int psh_gen_sky(const psh_sky_spec *g, psh_data *d_out, double *d_nz, size_t n);

/* ---- context layer -------------------------------------------------------- */

typedef struct psh_ctx psh_ctx;

/* Create a context for one configuration on the current HIP device.
 * nmock_max: number of independent mocks a batched call may hold (>= 1). */
int psh_ctx_create(const psh_conf *conf, int nmock_max, psh_ctx **out);
int psh_ctx_destroy(psh_ctx *ctx);

/* Fix the box (size and lower corner) explicitly, e.g. for simulation boxes
 * or precomputed survey boxes; computes the k-bin geometry. */
int psh_ctx_set_box(psh_ctx *ctx, const double bsize[3], const double bmin[3]);

/* Number of k bins of the current box (after psh_ctx_set_box or a run). */
int psh_ctx_nbin(const psh_ctx *ctx);

/* Run the whole path on host catalogues (same semantics as
 * psh_genr_mesh + psh_powspec, but the caller keeps ownership of the
 * particles, which are copied to the device and not modified).
 * out arrays must hold psh_ctx_nbin() bins (call after the box is known:
 * for surveys without BOX_SIZE use psh_ctx_define_box first). */
int psh_ctx_run_host(psh_ctx *ctx, psh_cata *cat, psh_pk *out);

/* Define the box from host catalogues (src/genr_mesh.c:509-578) without
 * running the path. */
int psh_ctx_define_box(psh_ctx *ctx, const psh_cata *cat);

/* Device-resident simulation-box run: d_data points to n psh_data records in
 * device memory (coordinates in [0, BOX_SIZE)); weight_sum = sum of w.
 * Auto power spectrum of one catalogue.  The box must be set. */
int psh_ctx_run_device(psh_ctx *ctx, const psh_data *d_data, size_t n,
    double weight_sum, psh_pk *out);

/* Batched mocks: nmock catalogues concatenated in d_data (device) with
 * offsets[nmock+1] (host) and weight_sum[nmock] (host); out[nmock].
 * graph != 0 replays the launch sequence as a captured hipGraph.  The
 * graph is recaptured when the batch's shape changes, and dropped whenever
 * the context reallocates a buffer it holds (psh_ctx_set_box,
 * psh_ctx_set_tiles, larger scratch for a later run). */
int psh_ctx_run_batch(psh_ctx *ctx, const psh_data *d_data,
    const size_t *offsets, const double *weight_sum, int nmock, int graph,
    psh_pk *out);

/* A survey (one data/random pair) with both catalogues in device memory:
 * box from their coordinate range (src/genr_mesh.c:427-578), then the
 * survey estimator of psh_powspec.  alpha, shot and norm are the
 * catalogue sums of src/read_cata.c:165-183 (see psh_cata). */
int psh_ctx_run_survey_device(psh_ctx *ctx, const psh_data *d_data, size_t nd,
    const psh_data *d_rand, size_t nr, double alpha, double shot, double norm,
    psh_pk *out);

/* Per-stage device time of the last run, milliseconds, measured with HIP
 * events on the context's stream (profiling on): [0] assignment incl. the
 * coordinate-bounds pass, [1] z and y FFT passes, [2] x FFT pass fused with
 * interlacing and binning, [3] total, [4] deposit kernel, [5] z passes,
 * [6] y passes, [7] bucketing kernels (count, scan, fill).  Survey runs:
 * [1] and [2] are the l = 0 field's passes, [5] the sum of the l > 0
 * fields' fused x passes (one per (l, m) and bin window), [6] the sum of
 * their z and y passes; [3] the survey estimator (the last run when there
 * are several catalogues). */
int psh_ctx_timings(const psh_ctx *ctx, double ms[8]);

/* Enable/disable per-stage event timing (adds event records between
 * stages; off by default). */
int psh_ctx_set_profiling(psh_ctx *ctx, int on);

/* Text of the last error on this context (or of the last drop-in call when
 * ctx is NULL). */
const char *psh_last_error(const psh_ctx *ctx);

/* Fault injection for tests of the error paths (no counterpart in the
 * reference, which exit()s on allocation failure inside parallel regions,
 * src/multipole.c:121-126).  PSH_FAULT_ALLOC: value n > 0 makes the n-th
 * device allocation of a context from now on fail (PSH_ERR_MEMORY), n < 0
 * every one, 0 none.  PSH_FAULT_GRAPH: value != 0 makes the instantiation of
 * a batch graph fail (PSH_ERR_DEVICE).  The environment variables
 * PSH_FAULT_ALLOC and PSH_FAULT_GRAPH set the same at load.  Library calls
 * return the error code and never exit. */
#define PSH_FAULT_ALLOC 1
#define PSH_FAULT_GRAPH 2
int psh_fault_inject(int what, long value);

/* ---- allocation-free run ---------------------------------------------------- */

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc -- the simulation-box run of psh_ctx_run_device (one
 * catalogue of n device records in [0, BOX_SIZE), auto power spectrum,
 * weight_sum = sum of w) without a context: every intermediate buffer
 * (meshes, tables, tile records, partial sums) is carved from one file-scope
 * device array of PSH_STATIC_SCRATCH_BYTES that each GPU the library runs on
 * holds until the process ends.  The call makes no device allocation (no
 * hipMalloc, hipFree or allocating library routine) and creates no stream:
 * it runs on the null stream of the current device and returns when the
 * results are on the host.  On entry out->nbin is the capacity of out's
 * arrays in bins (PSH_ERR_ARG when the run has more; psh_ctx_nbin of a
 * context with the same conf gives the count).  A workload whose buffers
 * exceed the array returns PSH_ERR_MEMORY with the size in the message (at
 * GRID_SIZE 128 interlaced, about 1e6 particles fit).  Calls on one GPU
 * share the array and the last call's contents stay in it: calls that
 * overlap on one GPU (other host threads) race. */
#define PSH_STATIC_SCRATCH_BYTES ((size_t)256 << 20)
int psh_run_device_static(const psh_conf *conf, const psh_data *d_data, size_t n,
    double weight_sum, psh_pk *out);

/* ---- device helpers (for hosts without their own HIP runtime bindings) ---- */
int psh_device_count(void);
int psh_device_malloc(void **ptr, size_t bytes);
int psh_device_free(void *ptr);
int psh_memcpy_h2d(void *dst, const void *src, size_t bytes);
int psh_memcpy_d2h(void *dst, const void *src, size_t bytes);
int psh_memcpy_d2d(void *dst, const void *src, size_t bytes);
int psh_device_synchronize(void);
/* Name of the device architecture (e.g. "gfx950"). */
int psh_device_arch(char *buf, int len);

/* ---- learned alias correction (no counterpart in the reference) ------------ */
/* A perceptron with inputs padded to 32, three hidden layers of 64 (ReLU) and
 * one output, trained and evaluated with bf16 matrix-core products and FP32
 * accumulation.  Parameters: one FP32 device array of psh_mlp_param_count()
 * values, W1[32][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][16]
 * b4[16] (only output column 0 is used; padding entries stay zero when
 * initialised to zero).  Samples: X[B][32] FP32 rows, targets T[B]. */
int psh_mlp_param_count(void);
/* Y[s] = network output of row s. */
int psh_mlp_forward(const float *d_params, const float *d_X, int B, float *d_Y);
/* Gradient of loss = mean (y - t)^2 over the B samples into d_grads
 * (overwritten); *loss (host, may be NULL) receives the loss. */
int psh_mlp_grad(const float *d_params, const float *d_X, const float *d_T, int B,
    float *d_grads, float *loss);
/* Adam (beta 0.9, 0.999) over `epochs` passes of contiguous mini-batches of
 * `batch` samples; *loss_out receives the final loss over all samples.
 * Gradients are summed in 64-bit fixed point, so equal inputs give equal
 * parameters on every run (psh_mlp_grad likewise). */
int psh_mlp_train(float *d_params, const float *d_X, const float *d_T, int B, int batch,
    int epochs, float lr, unsigned seed, float *loss_out);

/* Library stage of the learned correction (no counterpart in the reference,
 * whose fallback is the Jing correction of src/multipole.c:71-99): boxes
 * binned in |mu| cells, cell c of a k bin holding the modes with |mu| in
 * [c / nmu, (c + 1) / nmu) (1 <= nmu <= 64; 1: no cells, the default).  Box
 * runs then keep, besides P_l, the raw sums S[l][b][c] = sum over the
 * cell's modes of |delta|^2 alias mult L_l(mu). */
int psh_ctx_set_mu_cells(psh_ctx *ctx, int nmu);
/* The raw cell sums of the last box run, [npole][nbin][nmu]. */
int psh_ctx_cells(const psh_ctx *ctx, double *cells);
/* The network's input rows of the last box run, [nbin * nmu][32] floats
 * (row b nmu + c; zero rows for bins without modes or power): k_b / k_Ny,
 * the cell's |mu| centre, the one-hot scheme, 16 band averages of
 * log10(un-subtracted P_0 / shot) over [k_Ny / 4, k_Ny], the bin's own
 * log10(un-subtracted P_0 / shot) and log10(nmu S_0[b][c] / sum_c S_0[b][c])
 * clipped to [-2, 2]. */
int psh_ctx_alias_rows(psh_ctx *ctx, float *X);
/* Sets (params != NULL) or clears the learned correction of a plain
 * (non-interlaced) box context with a monopole and |mu| cells: nnet
 * networks, host params [nnet][psh_mlp_param_count()].  Its box runs
 * (psh_ctx_run_device, psh_ctx_run_host and the drop-in layer, auto
 * spectra) then compute on the device the rows above, the networks'
 * outputs f_n and the corrected sums sum_c exp(mean_n f_n[b][c]) S[l][b][c]
 * -- an estimate of the interlaced estimator's sums -- before the usual
 * normalisation, so the shifted grid's mesh and FFTs are skipped.  The
 * networks are trained on log(S_il / S_plain) of the monopole cell sums of
 * interlaced and plain runs (powspec_amd/mlp.py); the interlaced estimator
 * stays the fallback and the judge of the error. */
int psh_ctx_set_alias_mlp(psh_ctx *ctx, const float *params, int nnet);

/* ---- tuning ----------------------------------------------------------------- */
/* No counterpart in the reference (its FFTW plans use FFTW_ESTIMATE): the
 * deposit tile shape sets the bucketing duplication, the deposit's LDS
 * footprint and which bucketing kernel runs.  Results depend on it only
 * through each tile's fixed-point scale (integer accumulation; differences
 * at the 1e-9 level relative to a tile's largest cell). */

/* Tiles of a context: cells per axis, powers of two, each at least the
 * window width (scheme + 1), z >= 4, and both grids' LDS tiles within
 * 160 KB (z rows padded by the window width on each side; 4-byte cells for
 * boxes, 8-byte for surveys).  Overrides the default 8x16x64 and any cache.
 * PSH_ERR_ARG when invalid. */
int psh_ctx_set_tiles(psh_ctx *ctx, const int tile[3]);

/* The tile of whole z rows with which the deposit also performs the z
 * transform (no counterpart in the reference): for a simulation box of
 * GRID_SIZE 128, 256, 512 or 1024, tile[0..2] with tile[2] = GRID_SIZE; a
 * context given these tiles (psh_ctx_set_tiles, a tuning cache) writes
 * half-complex rows from the deposit and skips the separate z pass.  Returns
 * 1, or 0 (tile untouched) for other sizes.  Host-only. */
int psh_fused_tile(int gsize, int intlace, int tile[3]);

/* The deposit tiles in use (each axis clipped to GRID_SIZE): the shape set by
 * psh_ctx_set_tiles, a tuning cache, PSH_TILE or the default.  PSH_ERR_ARG
 * before the context has a box (surveys: before their first run). */
int psh_ctx_get_tiles(const psh_ctx *ctx, int tile[3]);

/* The box of the context's grid: side lengths and lower corner (surveys: the
 * padded box of the last run's catalogues, src/genr_mesh.c:427-578, which the
 * reference reports in its log).  PSH_ERR_ARG before the context has a box. */
int psh_ctx_get_box(const psh_ctx *ctx, double bsize[3], double bmin[3]);

/* Applies the fastest entries of a tuning cache for the context's device
 * architecture, grid size, scheme, interlacing and batch size (rounded down
 * to a power of two): deposit tiles and the y-pass schedule.  PSH_ERR_FILE /
 * PSH_ERR_ARG when none applies. */
int psh_ctx_load_tuning(psh_ctx *ctx, const char *cache_path);

/* FFT schedules (no counterpart in the reference, whose FFTW plans use
 * FFTW_ESTIMATE, src/fftw_define.h:64): the column (y) pass has schedules
 * 0..5 -- columns per tile (16, 16, 16, 16, 32, 8), values per thread
 * (32, 32, 32, 16, 32, 16) and radix plan (0: fewest passes of radix <= 16,
 * larger first; 1: the same passes, smaller first; 2: one pass more) -- of
 * which those whose LDS fits and that differ from schedule 0 apply at a
 * grid size.  Results differ between schedules by FP32 rounding only.
 * psh_ctx_set_fft_schedule refuses (PSH_ERR_ARG) one that does not apply. */
int psh_ctx_set_fft_schedule(psh_ctx *ctx, int ysched);
int psh_ctx_get_fft_schedule(const psh_ctx *ctx, int *ysched);

/* The radix plan of a length-n transform (no counterpart in the
 * reference, whose FFTW plans are opaque), n = 2^a, 3 * 2^a or 5 * 2^a
 * (a >= 1) up to 4096 -- the lengths of the x and y passes (GRID_SIZE) and
 * of the z pass's complex rows (GRID_SIZE / 2): plan 0, 1 or 2 as above;
 * writes the radix of each pass in order into radices[0..max) and returns
 * the number of passes, 0 for any other n.  Host-only (no device call);
 * used to check the plans on the CPU. */
int psh_fft_plan(int n, int plan, int *radices, int max);

typedef struct {
  int ysched;      /* fastest y-pass schedule               */
  double ms;       /* its median time per launch            */
  int ncand;       /* schedules timed                       */
  double ms_default; /* schedule 0's median time            */
} psh_tune_fft;

/* Times every applicable y-pass schedule on zeroed meshes of conf's grid
 * size (nmock meshes per launch, the live-column pruning of conf's
 * binning), appends "fft arch Ng intlace nmock_class y ysched ms" to
 * cache_path (when not NULL; psh_ctx_load_tuning applies the fastest entry
 * for the context) and returns the winner. */
int psh_autotune_fft(const psh_conf *conf, int nmock, const char *cache_path,
    psh_tune_fft *out);

typedef struct {
  int tile[3];     /* fastest shape                          */
  double ms;       /* its median time per run                */
  int ncand;       /* candidates timed                       */
} psh_tune;

/* Times candidate tile shapes (2^11..2^14 cells) for the grid size, scheme
 * and interlacing of conf on nmock mocks (batched when > 1): the caller's
 * device catalogues d_data with offsets[nmock + 1] (unit weights assumed for
 * the timing runs), or, when d_data is NULL, a uniform box of npart
 * particles per mock generated on the device.  Appends "arch Ng scheme
 * intlace nmock npart tx ty tz ms" to cache_path (when not NULL) and returns
 * the winner. */
int psh_autotune(const psh_conf *conf, const psh_data *d_data, const size_t *offsets,
    size_t npart, int nmock, const char *cache_path, psh_tune *out);

/* ---- catalogue ingest ----------------------------------------------------
 * Replaces the catalogue reader on the way into genr_mesh: read_cata
 * (src/read_cata.h:85, src/read_cata.c:100-183) -> read_ascii_data
 * (io/read_file.h, io/read_ascii.c:533-1142), whose per-object weight rule
 * and sums (io/read_ascii.c:875-903) the ingest applies on the device:
 *   survey: w = w_comp * w_fkp, sums w_comp, w^2 and w_comp w_fkp^2 n(z)
 *           (wdata/wrand, sumw2, sumw2n of read_cata);
 *   box:    w = w_comp, sum w_comp.
 * Absent w_comp and w_fkp columns read as 1, an absent n(z) as 0.  Columns
 * are float32 or float64 (all alike); the output is device psh_data.
 * The source is a binary SoA catalogue file (this library's format below)
 * or the caller's host columns.  Reading, the DMA over PCIe and the
 * conversion are pipelined through two pinned buffers on the ingest's own
 * stream and host thread: psh_ingest_file / psh_ingest_host return at once
 * and psh_ingest_wait completes, so the caller's context can work on the
 * previous catalogue meanwhile.  ASCII catalogues, with column numbers or
 * with libast expressions and a selection, go through psh_ingest_ascii and
 * psh_ingest_ascii_expr below. */
#define PSH_COL_X 1u
#define PSH_COL_Y 2u
#define PSH_COL_Z 4u
#define PSH_COL_WCOMP 8u
#define PSH_COL_WFKP 16u
#define PSH_COL_NZ 32u
#define PSH_CAT_MAGIC "PSHCAT01"

/* Binary catalogue file: this 64-byte header, then each present column
 * (bit order x, y, z, w_comp, w_fkp, nz) as n contiguous values of `dtype`
 * bytes, little endian. */
typedef struct {
  char magic[8];          /* PSH_CAT_MAGIC (8 bytes, no terminator)    */
  uint64_t n;             /* objects                                   */
  uint32_t dtype;         /* 4 (float32) or 8 (float64)                */
  uint32_t cols;          /* PSH_COL_* mask; x, y and z are required   */
  uint8_t reserved[40];
} psh_cat_header;

typedef struct {
  uint64_t n;             /* objects written to the device             */
  double sumw;            /* sum of w_comp                             */
  double sumw2;           /* sum of (w_comp w_fkp)^2 (surveys)         */
  double sumw2n;          /* sum of w_comp w_fkp^2 n(z) (surveys)      */
  double bytes;           /* column bytes moved host -> device         */
  double seconds_read;    /* host time spent reading/copying chunks    */
  double seconds;         /* wall time of the whole ingest             */
} psh_ingest_result;

typedef struct psh_ingest psh_ingest;

/* Pinned staging for chunks of `chunk_objects` objects (0: 2^21) on the
 * current device. */
int psh_ingest_create(size_t chunk_objects, psh_ingest **out);
int psh_ingest_destroy(psh_ingest *ing);

/* Write columns col_ptrs[c] (for each bit c of cols) as a binary catalogue. */
int psh_cat_write(const char *path, uint64_t n, int dtype, unsigned cols,
    const void *const col_ptrs[6]);

/* Start ingesting a binary catalogue file / host columns into d_out (device,
 * room for cap objects).  sim != 0 selects the box weight rule.  One ingest
 * in flight per object; finish it with psh_ingest_wait. */
int psh_ingest_file(psh_ingest *ing, const char *path, int sim, psh_data *d_out,
    size_t cap);
int psh_ingest_host(psh_ingest *ing, uint64_t n, int dtype, unsigned cols,
    const void *const col_ptrs[6], int sim, psh_data *d_out, size_t cap);

/* ASCII catalogues as the reference reads them (io/read_ascii.c:533-1142,
 * CONF keys DATA_SKIP, DATA_COMMENT, DATA_FORMATTER, DATA_POSITION,
 * DATA_WT_COMP, DATA_WT_FKP, DATA_NZ), parsed on the device: the first
 * `skip` lines are dropped, then lines whose first non-blank character is
 * `comment` (0: none) and blank lines; each remaining line is read with the
 * formatter's conversions as whitespace-separated tokens.  Conversions:
 * %lf %le %lg %lE %lG %lF (values), %d %i %ld %li %u %lu (integer values),
 * %*<conversion> (skipped; numeric ones must parse) and %*s %*c %*[...]
 * (any token).  Columns are 1-based among the conversions that are not
 * suppressed, like `$n` in the reference's expressions (psh_ingest_ascii_expr
 * below takes whole expressions and a selection).  Numbers
 * are converted exactly (the correctly rounded double, as strtod); a line
 * that does not match stops the ingest with PSH_ERR_FILE naming its line. */
typedef struct {
  long skip;              /* header lines                               */
  char comment;           /* comment character, 0: none                 */
  const char *fmtr;       /* e.g. "%lf %lf %lf %*s %lf"                 */
  int pos[3];             /* columns of x, y, z (RA, Dec, z for surveys) */
  int wcomp, wfkp, nz;    /* columns, 0: absent                          */
} psh_ascii_fmt;

int psh_ingest_ascii(psh_ingest *ing, const char *path, const psh_ascii_fmt *fmt,
    int sim, psh_data *d_out, size_t cap);

/* The same with the reference's libast expressions in place of column
 * numbers (CONF keys DATA_POSITION, DATA_WT_COMP, DATA_WT_FKP, DATA_NZ and
 * DATA_SELECTION; io/read_ascii.c:656-735 builds them, :857-903 evaluates
 * them per line; the column numbers of `fmt` are ignored).  Value
 * expressions are evaluated in double: numbers, $1..$9 and ${n} (1-based
 * among the value conversions), + - * / % (fmod) ** (pow, left-associative
 * like every binary operator), unary -, abs( sqrt( ln( log( (base 10) and
 * parentheses.  The selection (NULL: keep every line) is a boolean
 * expression: comparisons < <= > >= == != of numbers, ! && || and
 * isfinite(; %d columns and numbers written with digits only are integers
 * there: two integers divide with truncation, take C's remainder and
 * libast's integer power, and the bitwise operators << >> & ^ | ~ take
 * integers (exact below 2^53).  Refused with PSH_ERR_ARG when compiled:
 * string literals, a boolean compared with a number, bitwise operators on
 * non-integers, more than 16 distinct columns, 64 operations or 16 nested
 * values per expression.  A line the selection rejects is dropped. */
typedef struct {
  const char *pos[3];            /* x, y, z (RA, Dec, z for surveys)       */
  const char *wcomp, *wfkp, *nz; /* NULL: absent                           */
  const char *sel;               /* NULL: no selection                     */
} psh_ascii_expr;

int psh_ingest_ascii_expr(psh_ingest *ing, const char *path, const psh_ascii_fmt *fmt,
    const psh_ascii_expr *ex, int sim, psh_data *d_out, size_t cap);

/* The ASCII ingest's decimal conversion on the host: the correctly rounded
 * double of the whole token s[0..len) (PSH_ERR_ARG when it is not a
 * supported decimal number). */
int psh_parse_double(const char *s, size_t len, double *out);

/* Wait for the ingest in flight; fills *res (may be NULL). */
int psh_ingest_wait(psh_ingest *ing, psh_ingest_result *res);

/* Text of the last ingest error. */
const char *psh_ingest_error(const psh_ingest *ing);

#ifdef __cplusplus
}
#endif
#endif
