/*
 * psoracle.h -- FP64 CPU oracle for the mesh -> FFT -> P_l(k) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch C99 restatement of the
 * reference's algorithm (cheng-zhao/powspec), used as the parity checker for
 * the HIP library and as the timed CPU baseline.  Only tests/, the smoke()
 * entry and bench.py's cpu_baseline leg may load it; the product library never
 * links, calls or falls back to it.
 *
 * Every function cites the reference file:line it follows.  Quirks of the
 * reference are reproduced on purpose (see the per-function notes).
 *
 * Parity status of each function (how it is pinned, see tests/):
 *   pso_assign         brute-force window sums, mass conservation, textbook
 *                      window shapes, translation by whole cells
 *   pso_def_box        closed form (hand-computed boxes)
 *   pso_alias_corr     closed form incl. the pik<=0.01 Taylor branch
 *   pso_kbins          closed form of nbin/kedge
 *   pso_precomp        mode-count invariant (weighted count == Ng^3),
 *                      numpy brute force of the bin of every mode
 *   pso_fft_r2c/c2r    numpy.fft.rfftn / irfftn, direct DFT, Parseval
 *   pso_interlace      half-cell-shifted delta pair (analytic phase)
 *   pso_count_*        constant-|F|^2 field closed form, numpy brute force
 *   pso_ylm            reference header compiled as-is (oracle/_ref),
 *                      addition theorem, scipy real-form combinations
 *   pso_legpoly        reference header compiled as-is, scipy eval_legendre
 *   pso_run            composition of the above; physics sanity (uniform
 *                      box -> P_0 ~ 0 after shot subtraction)
 *   pso_legauss        numpy.polynomial.legendre.leggauss
 *   pso_cmvdist(_order) closed forms: Einstein-de Sitter and de Sitter
 *                      distances; convergence of the order
 *   pso_spline_*       scipy CubicSpline(bc_type='natural'); exact on
 *                      linear samples; out-of-range rule
 * FFTW3 (the reference's external FFT) is absent from this image, so the
 * reference itself is unbuildable here; the FFT is pinned against the DFT
 * definition instead.
 */
#ifndef PSORACLE_H
#define PSORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSO_ASSIGN_NGP 0
#define PSO_ASSIGN_CIC 1
#define PSO_ASSIGN_TSC 2
#define PSO_ASSIGN_PCS 3

#define PSO_MAX_ELL 6

/* Error codes, same values as the reference (src/define.h:112-124). */
#define PSO_ERR_MEMORY (-1)
#define PSO_ERR_ARG    (-2)
#define PSO_ERR_CATA   (-11)
#define PSO_ERR_MESH   (-13)
#define PSO_ERR_PK     (-14)

/* The CONF fields read on the path (src/load_conf.h:40-93), with kmin/kmax
 * already log10-converted when logscale is set (src/load_conf.c:1351-1355). */
typedef struct {
  int ndata;            /* number of data catalogues: 1 or 2      */
  int issim;            /* CUBIC_SIM                               */
  int gsize;            /* GRID_SIZE                               */
  int assign;           /* PARTICLE_ASSIGN                         */
  int intlace;          /* GRID_INTERLACE                          */
  int npole;            /* number of multipoles                    */
  int poles[PSO_MAX_ELL + 1]; /* sorted, unique                    */
  int has_bsize;        /* BOX_SIZE given (reference: bsize!=NULL) */
  double bsize[3];      /* BOX_SIZE                                */
  double bpad[3];       /* BOX_PAD                                 */
  double los[3];        /* LINE_OF_SIGHT                           */
  int logscale;         /* LOG_SCALE                               */
  double kmin, kmax, kbin;
  int isauto[2];
  int iscross;
} pso_conf;

/* Catalogues: AoS records (x, y, z, w), the reference's DATA
 * (src/read_cata.h:42-45), plus the per-catalogue sums of CATA
 * (src/read_cata.h:48-59).  alpha/shot/norm are inputs for surveys and are
 * written for simulation boxes (src/genr_mesh.c:904-910). */
typedef struct {
  int num;
  double *data[2];
  size_t ndata[2];
  double *rand[2];
  size_t nrand[2];
  double wdata[2], wrand[2];
  double alpha[2], shot[2], norm[2];
} pso_cata;

/* Output, equivalent to the reference's PK (src/multipole.h:38-61).  All
 * arrays are caller-allocated with room for `cap` bins. */
typedef struct {
  int cap;              /* capacity in bins                        */
  int nbin;             /* written: number of bins                 */
  int nl;               /* written: number of multipoles           */
  double *k;            /* [cap]                                   */
  double *kedge;        /* [cap + 1]                               */
  double *km;           /* [cap]                                   */
  uint64_t *cnt;        /* [cap]                                   */
  double *lcnt;         /* [nl * cap] (simulation boxes)           */
  double *pl[2];        /* [nl * cap] auto spectra, l-major        */
  double *xpl;          /* [nl * cap] cross spectrum               */
  double bsize[3];      /* written: box size                       */
  double bmin[3];       /* written: lower box corner               */
  double cmax[3];       /* written: upper coordinate bound         */
} pso_result;

/* ---- building blocks ---------------------------------------------------- */

/* src/genr_mesh.c:50-412 (ngp/cic/tsc/pcs).  data: n AoS records. */
void pso_assign(int scheme, const double *data, size_t n, int Ng,
    const double bmin[3], const double L[3], double *rho);

/* src/genr_mesh.c:427-492 (get_coord_bound), accumulating into min/max. */
void pso_coord_bound(const double *data, size_t n, double min[3],
    double max[3]);

/* src/genr_mesh.c:509-578 (def_box).  Returns 0 or PSO_ERR_CATA. */
int pso_def_box(const double min[3], const double max[3], int issim,
    const double *bsize, const double bpad[3], double size[3],
    double bmin[3]);

/* src/genr_mesh.c:590-602 (shift_cat). */
void pso_shift_cat(double *data, size_t n, const double smin[3],
    const double L[3]);

/* src/multipole.c:46-100 (alias_corr). */
double pso_alias_corr(int intlace, int assign, double pik);

/* src/multipole.c:334-394 (powspec_init bin set-up).  Returns nbin (>=1) or
 * a negative error code; writes nbin+1 edges and nbin centres when the
 * pointers are non-NULL and cap suffices. */
int pso_kbins(int Ng, const double L[3], int logscale, double kmin,
    double kmax, double dk, int cap, double *kedge, double *kcen);

/* src/multipole.c:111-257 (powspec_precomp).  alias: Ng*Ng*(Ng/2+1). */
void pso_precomp(int Ng, const double L[3], int intlace, int assign,
    int logscale, int nbin, const double *kedge, double dk, double *alias,
    uint64_t *cnt, double *km);

/* Unnormalised 3-D DFT, sign -1, half-complex output [Ng][Ng][Ng/2+1]
 * interleaved (re,im): FFTW r2c semantics (src/fftw_define.h:47-64). */
void pso_fft_r2c(int Ng, const double *in, double *out);
/* Unnormalised inverse (sign +1), FFTW c2r semantics; `in` is preserved. */
void pso_fft_c2r(int Ng, const double *in, double *out);

/* src/multipole.c:461-487 (interlace combine, in place on F0). */
void pso_interlace(int Ng, double *F0, const double *F1);

/* src/multipole.c:778-898 / 911-1031 (count_mode_sim_lin/log).
 * pl, lcnt: [nl*nbin], accumulated into (lcnt is zeroed first). */
void pso_count_sim(int Ng, const double L[3], int logscale, int nbin,
    const double *kedge, double dk, const double los[3], int nl,
    const int *poles, const double *alias, const double *F1,
    const double *F2, double *pl, double *lcnt);

/* src/multipole.c:579-660 / 673-754 (count_mode_lin/log).  pl: [nbin]. */
void pso_count_survey(int Ng, const double L[3], int logscale, int nbin,
    const double *kedge, double dk, const double *alias, const double *F1,
    const double *F2, double *pl);

/* math/spherical.h:54-260 (YlmR_l1..l6) restated from the definition of the
 * real-form spherical harmonics; math/legpoly.h:47-61. */
double pso_ylm(int ell, int m, double cost, double sint, double cosp,
    double sinp);
double pso_legpoly(int ell, double x);

/* src/mp_template.c:48-149 (dens_k<ell>) for one catalogue.
 * Fr: real field [Ng^3]; smin_g: box corner in grid units
 * (src/genr_mesh.c:913-914); Fkl: output [Nc complex]; Frl, Fka scratch. */
void pso_dens_kl(int ell, int Ng, const double L[3], const double smin_g[3],
    const double *Fr, const double *alias, double *Frl, double *Fka,
    double *Fkl);

/* ---- the whole path ----------------------------------------------------- */

/* genr_mesh (src/genr_mesh.c:874-926) followed by powspec
 * (src/multipole.c:1179-1278).  Catalogue coordinates are shifted in place
 * like the reference does (src/genr_mesh.c:590-602).  Returns 0 or a
 * negative error code. */
int pso_run(const pso_conf *conf, pso_cata *cat, pso_result *res);

/* Stage timings of the last pso_run (seconds): assign, fft, bin, total. */
void pso_last_timings(double t[4]);

/* Number of OpenMP threads the oracle uses (1 without OpenMP). */
int pso_num_threads(void);

/* ---- survey coordinate conversion (src/cnvt_coord.c) -------------------- */

/* Gauss-Legendre rule of order n (4..32): the n/2 positive abscissas, their
 * weights and, for odd n, the weight of x = 0 (w0); written from the
 * Legendre polynomials (Newton iteration), not from a table. */
void pso_legauss(int n, double *x, double *w, double *w0);

/* Radial comoving distance [Mpc/h] by n-point Gauss-Legendre integration of
 * c/(100 E(z)) from 0 to z, E^2 = Om (1+z)^3 + Ok (1+z)^2 + OL (1+z)^widx
 * (src/cnvt_coord.c:241-292). */
double pso_cmvdist(double om, double ol, double ok, double widx, int n, double z);

/* Order chosen by the reference's convergence test on 128 redshifts in
 * [zmin, zmax] (src/cnvt_coord.c:138-239, 306-341); -1 on failure. */
int pso_cmvdist_order(double om, double ol, double ok, double widx, double err,
    double zmin, double zmax);

/* Natural cubic spline: second derivatives of the samples
 * (math/cspline.c cspline_ypp), and the interpolation with the reference's
 * binary search (src/cnvt_coord.c:97-111), HUGE_VAL outside [z0, z_last). */
int pso_spline_ypp(const double *x, const double *y, size_t n, double *ypp);
double pso_spline_dist(const double *zs, const double *ds, const double *ypp,
    size_t n, double z);

/* In-place (RA, Dec, z) -> (x, y, z) of n records (src/cnvt_coord.c:122-134,
 * 372-386): integration when zs == NULL, else spline interpolation. */
void pso_cnvt(double *d, size_t n, double om, double ol, double ok,
    double widx, int order, const double *zs, const double *ds,
    const double *ypp, size_t nsp);

#ifdef __cplusplus
}
#endif
#endif
