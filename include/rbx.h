/*
 * rbx.h -- C ABI of librbx.so: MI355X (gfx950) kernels for RadioBEAR's
 * per-layer absorption and per-ray brightness-temperature integration.
 *
 * Conventions
 *   - Every entry point returns an int status: RBX_OK (0) or a negative code;
 *     rbx_last_error() gives a thread-local message.  No C++ exception crosses
 *     the ABI.
 *   - Array arguments are DEVICE pointers (HBM), owned by the caller, dense
 *     and C-ordered as documented; the only host pointers are the small
 *     descriptor structs and the host columns given to rbx_catalog_set_lines.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream); all
 *     work is enqueued asynchronously on it.
 *   - The library allocates device memory only inside an rbx_catalog handle,
 *     which it frees in rbx_catalog_destroy; scratch memory a call needs (the
 *     rbx_alpha workspace) is passed in by the caller.
 *
 * Reference interfaces replaced (RadioBEAR 2.0.1, paths relative to the
 * reference repository):
 *   rbx_alpha           Alpha.get_layers -> get_single_layer -> get_alpha_from_calc
 *                       -> <constituent>.alpha -> total_layer_alpha
 *                       (radiobear/alpha.py:151-305; constituents/nh3/nh3_hs_sjs.py:6-26,
 *                        nh3/nh3_hs.py:96-340, nh3/nh3_sjs.py:26-128, h2o/h2o_bk.py:65-189,
 *                        h2s/h2s_ddb.py:42-87, ph3/ph3_jh.py:64-108, h2/h2_jj_ddb.py:7-38;
 *                        the other planets' and optional absorbers nh3/nh3_dbs_sjs.py:6-26,
 *                        nh3/nh3_dbs.py:66-313, co/co_ddb.py:22-97, clouds/clouds_idp.py:6-101,
 *                        h2/h2_orton.py:17-222 with h2/h2_jj.py:7-22)
 *   rbx_alpha_tangent   (no reference counterpart: d alpha / d{T, q_NH3, q_H2O} for the adjoint)
 *   rbx_catalog_*       the line-list loads readInputFiles of the same modules
 *                       (nh3_hs.py:88-93, nh3_sjs.py:17-23, h2s_ddb.py:14-39,
 *                        ph3_jh.py:18-61, h2o_bk.py:17-62, co_ddb.py:14-19,
 *                        h2_orton.py:17-97)
 *   rbx_raypaths        raypath.compute_ds / findEdge / Shape._calcEllipse
 *                       (radiobear/raypath.py:39-273, radiobear/shape.py:223-298)
 *   rbx_integrate       Brightness.single's tau / W / Tb loop and final spectrum
 *                       (radiobear/brightness.py:30-126)
 *   rbx_integrate_profiles  the .tau / .W / .Tb_lyr attributes Brightness.single keeps
 *                       (radiobear/brightness.py:53-120)
 *   rbx_jacobian        (no reference counterpart: dTb/dT and dTb/dq per layer)
 *   rbx_integrate_adjoint, rbx_alpha_vjp  (no reference counterpart: the vector-Jacobian
 *                       product of Brightness.single's Tb, brightness.py:93-117, with respect
 *                       to the layer alpha and T, and its chain through alpha.py:151-192 to
 *                       d/d{T, q_NH3, q_H2O} per layer)
 */
#ifndef RBX_H
#define RBX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBX_ABI_VERSION 10

/* status codes */
#define RBX_OK          0
#define RBX_ERR_ARG    -1   /* invalid argument (shape, null pointer, enum) */
#define RBX_ERR_HIP    -2   /* HIP runtime error */
#define RBX_ERR_NOMEM  -3   /* device allocation failed */
#define RBX_ERR_STATE  -4   /* catalogue lacks a block the selected absorbers need */

/* arithmetic modes */
#define RBX_FP64        0   /* plain FP64 everywhere (IEEE division, FP64 exp) */
#define RBX_FP32        1   /* FP32 line sums, compensated (FP64 outer accumulation);
                               FP64 detuning, geometry, half steps and tau prefix sum;
                               point rays: e^-tau as the FP32 exp2 of the FP32-rounded
                               exponent, weights and trapezoid sums in FP32 added into
                               FP64 totals every batch of steps; disc rays and batches:
                               e^-tau from an FP64 range reduction and an FP32 exp2 */

/* line-block kinds and their columns (float64, one value per line) */
#define RBX_LINES_NH3_INV  0  /* fo, Io, Eo, gammaNH3o                     (ammonia.npz) */
#define RBX_LINES_NH3_ROT  1  /* fo_rot, Io_rot, Eo_rot, gNH3_rot, gH2_rot, gHe_rot */
#define RBX_LINES_NH3_V2   2  /* fo_v2, Io_v2, Eo_v2 */
#define RBX_LINES_NH3_SJS  3  /* f0, I0, E, G0                             (nh3.npz) */
#define RBX_LINES_H2S      4  /* f0, I0, E, GH2S                           (h2s.npz) */
#define RBX_LINES_PH3      5  /* f0, I0, E, WgtI0, WgtFGB, WgtSB  (ph3jh.npz + PH3WGT.npz) */
#define RBX_LINES_H2O      6  /* f_o, I_o, E_o, w_s, x_s, w_h2, x_h2, w_he, x_he, SR
                                 (h2o_bk schema, widths in GHz/bar) */
#define RBX_LINES_CO       7  /* f0, I0, E                                 (co.npz) */
#define RBX_TABLE_H2_ORTON 8  /* Orton's H2 CIA tables; n = table frequencies, and the
                                 columns are NOT n long: 0 frequencies (GHz, n, ascending),
                                 1 natural-log absorption [6 n][RBX_ORTON_NTEMP] (tables
                                 eh2h2, nh2h2, eh2he, nh2he, eh2ch4, nh2ch4, one row per
                                 frequency), 2 the table temperatures (K, RBX_ORTON_NTEMP,
                                 ascending) -- orton_H2.tables as h2_orton.py:17-97 reads it */
#define RBX_NUM_LINE_KINDS 9
#define RBX_ORTON_NTEMP    10

/* constituent slots, in the reference's sorted order (alpha.py:80) */
#define RBX_C_H2   0
#define RBX_C_H2O  1
#define RBX_C_H2S  2
#define RBX_C_NH3  3
#define RBX_C_PH3  4
#define RBX_C_CO   5
#define RBX_C_CLOUDS 6
#define RBX_NUM_CONSTITUENTS 7   /* the total sums them by name: clouds, co, h2, h2o, h2s, nh3, ph3 */

/* rows of the mixing-ratio table mix[RBX_NUM_SPECIES][nl] */
#define RBX_X_H2   0
#define RBX_X_HE   1
#define RBX_X_CH4  2
#define RBX_X_NH3  3
#define RBX_X_H2O  4
#define RBX_X_H2S  5
#define RBX_X_PH3  6
#define RBX_X_CO   7   /* read by co_ddb only (its molar mass and refractivity are chemistry.py's NULL, 0) */
#define RBX_NUM_SPECIES 8

/* rows of the cloud-density table cloud[RBX_NUM_CLOUDS][nl] (g/cm^3, the reference's
 * atm.cloud rows H2O, SOLN, NH4SH, NH3, H2S, CH4), in the order clouds_idp sums them */
#define RBX_CL_H2O   0   /* water ice; other_dict flag ice_p */
#define RBX_CL_SOLN  1   /* liquid water / solution; water_p */
#define RBX_CL_NH4SH 2   /* nh4sh_p */
#define RBX_CL_NH3   3   /* ammonia ice; nh3ice_p */
#define RBX_CL_H2S   4   /* h2sice_p */
#define RBX_CL_CH4   5   /* the module's flag is 'ch4' (clouds_idp.py:46) */
#define RBX_NUM_CLOUDS 6

typedef struct rbx_catalog rbx_catalog;

/* Absorber selection: which formalism each constituent uses (0 = not loaded). */
typedef struct rbx_absorbers {
    int h2;    /* 1: h2_jj_ddb h2state 'e'; 2: h2_jj_ddb h2state 'n';
                  3: h2_orton h2state 'e'; 4: h2_orton h2state 'n' */
    int h2o;   /* 1: h2o_bk */
    int h2s;   /* 1: h2s_ddb */
    int nh3;   /* 1: nh3_hs_sjs; 2: nh3_hs; 3: nh3_sjs; 4: nh3_dbs_sjs; 5: nh3_dbs
                  (constituents/nh3/nh3_dbs_sjs.py:6-26, nh3_dbs.py:66-313; the Saturn,
                  Uranus and Neptune default, <planet>/config.par:23) */
    int ph3;   /* 1: ph3_jh */
    int co;    /* 1: co_ddb (the Neptune default) */
    int co_shape;     /* co_ddb coshape: 0 'voigt' (the default), 1 'vvw', 2 'diff' -- each the sum
                         of the line shapes without line strengths, as co_ddb.py:86-91 -- or
                         3 (any other value): pressure blend of the two, times the strengths */
    int clouds;       /* 1: clouds_idp */
    int cloud_types;  /* clouds_idp: bit RBX_CL_x set = that cloud's other_dict flag is > 0 */
} rbx_absorbers;

/* Planet shape and view for the ray tracer. */
typedef struct rbx_planet {
    double req;             /* km, equatorial radius at p_ref (config.Req) */
    double rpol;            /* km, polar radius at p_ref (config.Rpol) */
    double orientation[2];  /* deg: [position angle, sub-observer planetographic latitude] */
    int gtype;              /* 0: 'ellipse'; 1: 'sphere' / 'circle' */
    int limb;               /* 0: 'shape'; 1: 'sec' (raypath.py:218-219) */
} rbx_planet;

int rbx_abi_version(void);

/* Copy the calling thread's last error message (NUL-terminated) into buf. */
int rbx_last_error(char *buf, size_t len);

/* Number of HIP devices visible to the library (host call, no GPU work). */
int rbx_device_count(int *count);

/* Launch shape of one call (no reference counterpart; radiobear_amd.autotune fills it).
 * Every entry point that makes a launch choice takes a `const rbx_launch *launch`:
 * NULL, or a field left 0, means the library's built-in choice for that call's shape.
 * Nothing is kept between calls: the same arguments give the same plan, so a
 * workspace sized by rbx_alpha_workspace_size(..., launch, ...) fits rbx_alpha(...,
 * launch, ...) on any thread. */
typedef struct rbx_launch {
    int alpha_k;        /* frequencies per lane of the forward line sums: 1, 2 or 4 (built in:
                           2 from 256 frequencies up, else 1) */
    int alpha_series;   /* rbx_alpha: 1 sums every line at every frequency; 2 sums per wavefront
                           tile only the lines whose poles lie within a factor 3 (FP32) or 6.7
                           (FP64) of the tile's x = f^2 range and adds the others as 16-term power
                           series (at most 8e-7 / 8e-12 of a far line's value: FP32 tables within
                           3e-7 of FP64 as the direct sums, FP64 tables within 1e-11 of the direct
                           FP64 sums); built in: 2 for requests of 4 or more wavefront tiles while
                           the workspace (the tiles' moments and near-line masks: about 0.1 MB per
                           layer at 1024 frequencies and 25553 lines) stays under 4 GiB, else 1 */
    int alpha_split;    /* parts each layer's direct line sums are split into: 1..16 (built in: a
                           model of the grid's last dispatch wave) */
    int far_split;      /* parts each layer's lines are split into in the series' classification
                           pass: 1..16 (built in: about 4096 blocks) */
    int int_unroll;     /* path steps per load batch of the integration: 4 or 8 (built in: 8) */
    int int_map;        /* integration block order: 1 frequency group major (each XCD keeps a
                           slice of the table for every ray), 2 ray major (each XCD runs every
                           frequency group of its rays: their ds rows are read once per ray and
                           the whole table lives in each XCD's L2); built in: ray major for
                           tables up to 3.5 MiB */
    int tan_split;      /* parts each layer's lines are split into by rbx_alpha_tangent: 1..16
                           (built in: 1 when layers x 64-column tiles fill the device, else
                           enough parts for about 4 wavefronts per SIMD) */
    int adj_rays;       /* rays per wavefront of rbx_integrate_adjoint: 1, 2, 4 or 8 (built in: 4);
                           their terms are summed in registers before each atomic add */
    int adj_copies;     /* copies of the adjoint's accumulators the atomic adds are spread over:
                           1..64 (built in: one per 4 ray blocks of an atmosphere, at most 32) */
    int reserved[7];    /* must be 0 */
} rbx_launch;

/* ---- line catalogue (device-resident, owned by the library) ---- */
int rbx_catalog_create(rbx_catalog **out);
/* Replace block `kind` with n lines; host_columns[i] is a HOST array of n
 * doubles for column i of that kind (see RBX_LINES_*).  n = 0 clears it.  The
 * line lists the forward sums stage (NH3 inversion, rotational, v2, H2S, PH3,
 * H2O) are stored sorted by centre frequency (column 0, stable). */
int rbx_catalog_set_lines(rbx_catalog *cat, int kind, int n, const double *const *host_columns);
int rbx_catalog_lines(const rbx_catalog *cat, int kind, int *n);
int rbx_catalog_destroy(rbx_catalog *cat);

/* ---- absorption: alpha[nl][nf] in cm^-1 ----
 * P, T            [nl]   bar, K (device)
 * mix             [RBX_NUM_SPECIES][nl] mixing ratios (device)
 * cloud           [RBX_NUM_CLOUDS][nl] cloud densities, g/cm^3 (device); read only
 *                 when clouds_idp is selected, may be NULL otherwise
 * freq            [nf]   GHz (device), the request order
 * freq_nh3        [nf]   GHz (device) or NULL: the frequency each output column
 *                 uses for NH3 (nh3_hs returns its <= 30 GHz half first,
 *                 nh3_hs.py:96-114); its first nf_nh3_low entries must be <= 30 GHz
 *                 and the rest > 30 GHz.  NULL means freq (which must then be
 *                 ordered that way, e.g. ascending) with nf_nh3_low computed by the caller.
 * scale           [RBX_NUM_CONSTITUENTS][nl] per-constituent per-layer scale
 *                 (total_layer_alpha, alpha.py:151-192) or NULL for 1.0
 * alpha           [nl][nf] float64 (RBX_FP64) or float32 (RBX_FP32), layer-major
 * by_constituent  [RBX_NUM_CONSTITUENTS][nl][nf] float64 unscaled per-constituent
 *                 alpha, or NULL (rows of constituents that are not loaded are not written)
 * workspace       device scratch of at least rbx_alpha_workspace_size bytes (8-byte
 *                 aligned) for the partial line sums; it must not be used by another
 *                 call that runs at the same time (a stream-ordered allocation is right)
 */
int rbx_alpha(const rbx_catalog *cat, const rbx_absorbers *absorbers,
              const double *P, const double *T, const double *mix, const double *cloud, int nl,
              const double *freq, const double *freq_nh3, int nf, int nf_nh3_low,
              const double *scale, int precision, void *alpha, double *by_constituent,
              const rbx_launch *launch, void *workspace, size_t workspace_bytes, void *stream);

/* rbx_alpha for callers that cannot allocate (no counterpart in the reference: for code
 * under HIP graph capture, tight serving loops and harnesses that forbid hipMalloc).  Same
 * arguments and results as rbx_alpha without the workspace: the partial sums go to a static
 * device array of RBX_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until
 * the process ends.  The call allocates nothing.  It uses the plan rbx_alpha would use with
 * `launch`; when that needs more scratch than the maximum and launch->alpha_series is not 2,
 * it takes the direct line sums with the most parts (up to that plan's) that fit.  A call
 * that still does not fit returns RBX_ERR_ARG (the direct sums need 32 B per layer and
 * frequency plus 8 B per layer, per part).  Calls on one GPU share its array: calls that
 * overlap in time on one GPU (other streams, other threads) race and must not be made. */
#define RBX_NOALLOC_SCRATCH_BYTES (64u << 20)
int rbx_alpha_noalloc(const rbx_catalog *cat, const rbx_absorbers *absorbers,
                      const double *P, const double *T, const double *mix, const double *cloud, int nl,
                      const double *freq, const double *freq_nh3, int nf, int nf_nh3_low,
                      const double *scale, int precision, void *alpha, double *by_constituent,
                      const rbx_launch *launch, void *stream);

/* Workspace bytes rbx_alpha needs for this shape and launch (host call). */
int rbx_alpha_workspace_size(const rbx_catalog *cat, const rbx_absorbers *absorbers, int nl, int nf,
                             int nf_nh3_low, int precision, const rbx_launch *launch, size_t *bytes);

/* ---- reuse of a saved absorption table (Alpha get_alpha='memory'/'file',
 * alpha.py:110-149 and total_layer_alpha alpha.py:151-192) ----
 * absorb          [nl][nf][nc] float64 per-constituent absorption in the saved
 *                 column order (Alpha.alpha_data)
 * scale           [nc][nl] per-constituent per-layer factor, or NULL for 1.0
 * alpha           [nl][nf] float64 (RBX_FP64) or float32 (RBX_FP32): sum_c scale * absorb
 */
int rbx_alpha_rescale(const double *absorb, int nl, int nf, int nc, const double *scale, int precision,
                      void *alpha, void *stream);

/* ---- absorption and its tangents for the adjoint (one fused pass) ----
 * Same inputs as rbx_alpha, for the h2_jj_ddb, h2o_bk, h2s_ddb, nh3_* and ph3_jh
 * absorbers (a selection with co_ddb, clouds_idp or h2_orton is rejected).  precision RBX_FP64 (plain FP64) or RBX_FP32
 * (FP32 line sums as in rbx_alpha; tangents good to ~1e-6 relative).
 * alpha           [nl][nf] float64 alpha, or NULL
 * dalpha          [3][nl][nf] float64: d alpha / d T, d alpha / d q_NH3,
 *                 d alpha / d q_H2O (q = mixing ratio), same layout as alpha.
 * nl may be the total layer count of several atmospheres laid end to end
 * (layers are independent here); see the *_batch entry points below.
 */
int rbx_alpha_tangent(const rbx_catalog *cat, const rbx_absorbers *absorbers,
                      const double *P, const double *T, const double *mix, int nl,
                      const double *freq, const double *freq_nh3, int nf, int nf_nh3_low,
                      const double *scale, int precision, double *alpha, double *dalpha,
                      const rbx_launch *launch, void *workspace, size_t workspace_bytes,
                      void *stream);

/* Workspace bytes rbx_alpha_tangent needs for this shape and launch (host call): each
 * layer's dual-number coefficient derivers, 1168 bytes per layer, and, when each layer's lines
 * are summed in P > 1 parts (the built-in choice until layers x 64-column tiles fill the
 * device), 96 P bytes per layer and frequency more (plus 24 P per layer). */
int rbx_alpha_tangent_workspace_size(const rbx_catalog *cat, const rbx_absorbers *absorbers, int nl, int nf,
                                     int nf_nh3_low, int precision, const rbx_launch *launch, size_t *bytes);

/* ---- ray geometry (FP64) ----
 * R, N            [nl] layer radii (km) and refractive indices, top layer first
 * b               [nray][2] impact parameters (fraction of R[0])
 * ds              [nray][ld] step lengths; out_units 0: km (Ray.ds),
 *                 1: half step in cm ((ds*1e3/1e-2)/2, what rbx_integrate consumes)
 * nsteps          [nray] entries of ds used; -1 = off planet (Ray.ds None)
 * ld >= nl - 1.
 */
int rbx_raypaths(const rbx_planet *planet, const double *R, const double *N, int nl,
                 const double *b, int nray, int out_units, double *ds, int ld, int *nsteps,
                 void *stream);

/* ---- brightness integration ----
 * alpha           [nl][nf] float64 (alpha_precision RBX_FP64) or float32 (RBX_FP32)
 * T               [nl] K
 * hds, ld, nsteps from rbx_raypaths with out_units 1
 * mode            0: point ray, W = alpha e^-tau; 1: disc average, W = 2 alpha E2(tau)
 * precision       RBX_FP64 or RBX_FP32 (how e^-tau is evaluated; tau is FP64 in both)
 * tau_cut         stop a ray's frequency group once every tau > tau_cut (<= 0: never)
 * Tb              [nray][nf] K; intW [nray][nf] integrated weight or NULL
 */
int rbx_integrate(const void *alpha, int alpha_precision, const double *T, int nl, int nf,
                  const double *hds, int ld, const int *nsteps, int nray,
                  int mode, int precision, double tau_cut, double *Tb, double *intW,
                  const rbx_launch *launch, void *stream);

/* The same integration with the frequencies taken in the order of cols (no counterpart in
 * the reference; Brightness.single, brightness.py:30-126, integrates every frequency to the
 * end of the path).  A wavefront of 64 frequencies stops at the tau cut only once all of them
 * are past it, so frequencies that pass it at similar depths should share wavefronts.
 * rbx_integrate_order writes such an order, cols [nf] (column j of the reordered table is
 * request frequency cols[j]; a permutation), ranking the frequencies by the path step at which
 * the on-planet ray with the longest path passes tau_cut (stable; at most 16384 frequencies
 * and 131072 layers),
 * and the table in that order, alpha_cols [nl][nf] (alpha's precision).  rbx_integrate_cols
 * integrates alpha_cols and writes Tb and intW to column cols[j] of each row, so its outputs
 * are rbx_integrate's up to what the cut itself changes (e^-tau_cut of a ray's weight).
 * factors: NULL, or rbx_integrate_factors' table for the same T, hds, ld, nsteps and nray:
 * point rays (mode 0) in RBX_FP32 precision with an RBX_FP32 table then read each ray's
 * path factors from it, where without it every frequency group recomputes them (the
 * trapezoid factors into LDS); the outputs are bit-identical either way, and other modes and
 * precisions ignore it.  rbx_integrate_factors writes that table, factors [nray][ld] records
 * of 16 bytes (16 nray ld bytes, 16-byte aligned): for path step k of a ray of n = nsteps - 1
 * steps, the FP64 half step times log2(e) (the FP32 integration scans tau log2 e) and two
 * FP32 trapezoid factors of point k + 1, h_k + h_{k+1} and T_{k+1} (h_k + h_{k+1}) (h_n = 0;
 * brightness.py:99-101 regrouped by path point); record n flags a path holding a NaN.  It
 * reads no alpha, so it can run beside the absorption on another stream (no counterpart in
 * the reference). */
int rbx_integrate_order(const void *alpha, int alpha_precision, int nl, int nf, const double *hds, int ld,
                        const int *nsteps, int nray, double tau_cut, int *cols, void *alpha_cols, void *stream);
int rbx_integrate_cols(const void *alpha_cols, int alpha_precision, const int *cols, const double *T, int nl,
                       int nf, const double *hds, int ld, const int *nsteps, int nray, int mode, int precision,
                       double tau_cut, double *Tb, double *intW, const void *factors, const rbx_launch *launch,
                       void *stream);
int rbx_integrate_factors(const double *T, int nl, const double *hds, int ld, const int *nsteps, int nray,
                          void *factors, void *stream);

/* ---- per-point profiles (Brightness.single's .tau, .W, .Tb_lyr, brightness.py:53-120) ----
 * tau, W, Tb_lyr  [nray][nl][nf] float64 at path points 0..nsteps-1 (0 beyond);
 * same arguments and arithmetic as rbx_integrate, no tau cut.
 */
int rbx_integrate_profiles(const void *alpha, int alpha_precision, const double *T, int nl, int nf,
                           const double *hds, int ld, const int *nsteps, int nray, int mode,
                           int precision, double *tau, double *W, double *Tb_lyr, void *stream);

/* ---- Jacobian rows ----
 * dalpha          [3][nl][nf] from rbx_alpha_tangent
 * J               [nray][3][nl][nf] float32: dTb/dT[l], dTb/dq_NH3[l] and dTb/dq_H2O[l] of
 *                 row (ray, f); layers a ray does not reach (or past tau_cut) are 0
 * Tb              [nray][nf] float64 or NULL
 */
int rbx_jacobian(const void *alpha, int alpha_precision, const double *dalpha, const double *T,
                 int nl, int nf, const double *hds, int ld, const int *nsteps, int nray,
                 double tau_cut, float *J, double *Tb, void *stream);

/* ---- many atmospheres in one launch (no reference counterpart; the retrieval
 * batch: Planet.run over a list of perturbed atmospheres, planet_base.py:209-221)
 * Per-layer arrays (R, N, P, T, mix rows, alpha, dalpha rows) of natm atmospheres
 * are laid end to end: atmosphere a owns layers [layer_off[a], layer_off[a+1])
 * of nl_total (layer_off: device int[natm + 1]); rbx_alpha / rbx_alpha_tangent
 * take nl = nl_total as they are.  ray_atm [nray] (device) gives each ray's
 * atmosphere; order [nray] (device, or NULL for 0..nray-1) is the order rays are
 * scheduled in (longest paths first balances uneven atmospheres); results are
 * written at each ray's own index either way.  A ray whose atmosphere index or
 * offsets are out of range gets nsteps = -2 and Tb = NaN.
 */
int rbx_raypaths_batch(const rbx_planet *planet, const double *R, const double *N, int nl_total,
                       const int *layer_off, int natm, const double *b, int nray, const int *ray_atm,
                       int out_units, double *ds, int ld, int *nsteps, void *stream);
int rbx_integrate_batch(const void *alpha, int alpha_precision, const double *T, int nl_total, int nf,
                        const int *layer_off, int natm, const double *hds, int ld, const int *nsteps,
                        const int *ray_atm, const int *order, int nray, int mode, int precision,
                        double tau_cut, double *Tb, double *intW, const rbx_launch *launch,
                        void *stream);
/* J [nray][3][ldj][nf] float32, ldj >= the layer count of every atmosphere (rows
 * past a ray's atmosphere are 0); precision as rbx_integrate's (e^-tau evaluation). */
int rbx_jacobian_batch(const void *alpha, int alpha_precision, const double *dalpha, const double *T,
                       int nl_total, int nf, const int *layer_off, int natm, const double *hds, int ld,
                       const int *nsteps, const int *ray_atm, const int *order, int nray, int precision,
                       double tau_cut, float *J, int ldj, double *Tb, void *stream);

/* ---- adjoint of the integration: the vector-Jacobian product ----
 * For a loss L with upstream gradient gTb [nray][nf] = dL/dTb (float64, device):
 *   galpha [nl][nf]  dL/dalpha[l][f] = sum_ray gTb dTb/dalpha[l][f]
 *   gT     [nl]      the direct part of dL/dT[l]: sum_ray,f gTb dTb/dT_l through T_l W_l
 * (both overwritten; float64).  The atomic adds of many rays go first to copies of
 * the accumulators in the workspace (rbx_integrate_adjoint_workspace_size), summed at the end.  Same arithmetic and tau_cut as rbx_jacobian: a forward pass
 * per ray, then the derivatives along the path with e^-tau recomputed.  order [nray]
 * (device, or NULL for 0..nray-1) groups rays: launch.adj_rays consecutive entries of it
 * share a wavefront, whose rays should be of one atmosphere and of similar path length
 * (longest first) for speed; any order gives the same result up to the order of FP64
 * additions.  An entry gTb = 0 contributes nothing (also where Tb is NaN); a NaN Tb under
 * a non-zero gTb makes the gradient NaN.  Paths are held fixed (the ray geometry is not
 * differentiated).  The batch form takes several atmospheres as rbx_integrate_batch. */
int rbx_integrate_adjoint(const void *alpha, int alpha_precision, const double *T, int nl, int nf,
                          const double *hds, int ld, const int *nsteps, int nray, const int *order,
                          int precision, double tau_cut, const double *gTb, double *galpha, double *gT,
                          const rbx_launch *launch, void *workspace, size_t workspace_bytes, void *stream);
int rbx_integrate_adjoint_batch(const void *alpha, int alpha_precision, const double *T, int nl_total, int nf,
                                const int *layer_off, int natm, const double *hds, int ld, const int *nsteps,
                                const int *ray_atm, const int *order, int nray, int precision, double tau_cut,
                                const double *gTb, double *galpha, double *gT, const rbx_launch *launch,
                                void *workspace, size_t workspace_bytes, void *stream);
/* Workspace bytes of one adjoint call (host call): copies of the accumulators the atomic
 * adds are spread over, 8 (copies nl nf + copies ceil(nf/64) nl) bytes, or 0 (natm = 1 for
 * rbx_integrate_adjoint). */
int rbx_integrate_adjoint_workspace_size(int nl, int nf, int nray, int natm, const rbx_launch *launch,
                                         size_t *bytes);

/* Chain through the absorption: grad [3][nl] (float64, overwritten) =
 *   dL/dT[l]     = gT[l] + sum_f galpha[l][f] dalpha[0][l][f]
 *   dL/dq_NH3[l] = sum_f galpha[l][f] dalpha[1][l][f],  dL/dq_H2O[l] = sum_f galpha dalpha[2]
 * dalpha [3][nl][nf] from rbx_alpha_tangent; gT may be NULL (no direct part). */
int rbx_alpha_vjp(const double *dalpha, const double *galpha, const double *gT, int nl, int nf,
                  double *grad, void *stream);

/* ---- atmosphere construction (computeProp on the device, atm_base.py:120-194) ----
 * For natm atmospheres laid end to end (layer_off as above), each ordered top
 * (lowest P) to bottom, from P [nl_total] bar, T [nl_total] K and
 * mix [RBX_NUM_SPECIES][nl_total]:
 *   Z [nl_total] km: copied from Z_in, or (Z_in NULL) hydrostatic with the deepest
 *     layer at 0: dz = (H_i + H_{i+1})/2 ln(P_{i+1}/P_i), H = R T / (m g), g = gm / req^2
 *     (a T perturbation moves the shells, which the ray paths then follow);
 *   R = req + Z - Z(layer nearest p_ref), N = 1 + 1e-6 sum_s refr_s(T) q_s P 293/T.
 * gm in km^3/s^2 (config.GM_ref), req in km (config.Req), p_ref in bar.
 */
int rbx_atmosphere_batch(const double *P, const double *T, const double *mix, const double *Z_in,
                         int nl_total, const int *layer_off, int natm, double req, double gm,
                         double p_ref, double *Z, double *R, double *N, void *stream);

/* computeProp's whole property table (atm_base.py:120-194; replaces the property array
 * AtmosphereBase.computeProp builds), for the same atmospheres from their layers' P, T, mix
 * and Z (rbx_atmosphere_batch's Z, or the caller's): props [RBX_NUM_PROPS][nl_total] with the
 * rows in the reference's order (config LP, default_config.json):
 *   Z km, R km, P bar, GM km^3/s^2, AMU g/mol, REFR, N, H km, LAPSE K/km, LAPSEP K/bar,
 *   g km/s^2.
 * GM is the mass above each layer summed from the top (dM = 1e11 rho 4 pi R^2 |dz|, rho =
 * AMU P / (R_gas T)) and referred to gm_ref at the layer nearest p_ref; g = GM / R^2,
 * H = R_gas T / (g AMU) / 1000; LAPSE, LAPSEP are |dT/dz|, |dT/dP| to the layer above (0 at
 * the top; a repeated z or P gives inf or NaN there, as IEEE division does).
 */
#define RBX_NUM_PROPS   11
#define RBX_PROP_Z      0
#define RBX_PROP_R      1
#define RBX_PROP_P      2
#define RBX_PROP_GM     3
#define RBX_PROP_AMU    4
#define RBX_PROP_REFR   5
#define RBX_PROP_N      6
#define RBX_PROP_H      7
#define RBX_PROP_LAPSE  8
#define RBX_PROP_LAPSEP 9
#define RBX_PROP_G      10
int rbx_atmosphere_props(const double *P, const double *T, const double *mix, const double *Z,
                         int nl_total, const int *layer_off, int natm, double req, double gm_ref,
                         double p_ref, double *props, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* RBX_H */
