/*
 * relxill_gpu.h -- C ABI of librelxill_gpu.so, the MI355X (gfx950) implementation
 * of relxill's relativistic smearing path (relline, relline_lp, relconv,
 * relconv_lp, plus a batched API and a serving loop).
 *
 * Drop-in boundary.  The four XSPEC local-model entry points below have the
 * exact signature and parameter order of the functions the reference's wrapper
 * generator emits (src/create_wrapper_xspec.py:153-162, parameter order from
 * src/modelfiles/lmodel_relxill_public.dat:1-44).  A host that loads the
 * reference library's symbols `lmodrelline`, `lmodrellinelp`, `lmodrelconv`,
 * `lmodrelconvlp` binds these instead.  `spectrum`, `fluxError` and `init` are
 * ignored as in the reference.  On failure the flux is zeroed and the status is
 * available from rx_last_status()/rx_last_error() (the reference throws
 * ModelEvalFailed through its extern "C" frame, src/LocalModel.cpp:143-160).
 *
 * No torch types, no C++ types: plain pointers and sizes.  Thread safety: every
 * entry point may be called from several host threads; calls are serialised
 * per device except the server API, which batches concurrent requests.
 */
#ifndef RELXILL_GPU_H_
#define RELXILL_GPU_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define RX_OK 0
#define RX_ERR_PARAM 1        /* parameters rejected (check_parameter_bounds) */
#define RX_ERR_UNSUPPORTED 2  /* feature out of scope (extended jet) */
#define RX_ERR_TABLE 10       /* table interpolation failed (grid invariants) */
#define RX_ERR_INPUT 50       /* relconv input flux with sum <= 0 (src/LocalModel.cpp:84-86) */
#define RX_ERR_DEVICE 100     /* HIP error / no device */
#define RX_ERR_ARG 101        /* invalid argument (sizes, null pointers) */
#define RX_ERR_NOMEM 102

/* ---- model ids: the reference's MOD_TYPE_* values (src/common.h:85-95) ---- */
#define RX_MODEL_RELLINE 1
#define RX_MODEL_RELLINELP 2
#define RX_MODEL_RELCONV 11
#define RX_MODEL_RELCONVLP 12
#define RX_MODEL_RELXILL (-1) /* relxill: broken power law, one zone, xillver table */
#define RX_MODEL_RELXILLLP (-2) /* relxilllp: lamp post, 10 radial zones, xillver table */
/* project id: relxill-style convolution of a resident angle-resolved
 * reflection set (rx_reflset_create) with relconv's parameters */
#define RX_MODEL_RELXILL_CONV (-1001)
/* project ids: the nthcomp-primary relxill models (relxillCp, relxilllpCp;
 * the reference's MOD_TYPE_RELXILL / MOD_TYPE_RELXILLLP with
 * PRIM_SPEC_NTHCOMP, src/ModelDefinition.cpp:38-50, :88) */
#define RX_MODEL_RELXILLCP (-1010)
#define RX_MODEL_RELXILLLPCP (-1011)
/* project id: the blackbody-primary relxill model (relxillNS; the reference's
 * MOD_TYPE_RELXILL with PRIM_SPEC_BB, src/ModelDefinition.cpp:89) */
#define RX_MODEL_RELXILLNS (-1012)
/* the non-relativistic xillver models (the reference's MOD_TYPE_XILLVER,
 * MOD_TYPE_XILLVER_NTHCOMP, MOD_TYPE_XILLVERNS, src/common.h:91-111) */
#define RX_MODEL_XILLVER 0
#define RX_MODEL_XILLVERCP 100
#define RX_MODEL_XILLVERNS (-101)

/* ---- precision of the batched path ---- */
#define RX_FP64 0 /* plain form: FP64 throughout */
#define RX_FP32 1 /* FP32 integrand and Romberg; FP64 decisions, edges, sums over radii, norms */

/* ===================================================================== */
/* XSPEC local-model entry points (replace the generated lmod<name>,       */
/* src/create_wrapper_xspec.py:153-162 -> xspec_C_wrapper_eval_model,      */
/* src/LocalModel.cpp:143-160).                                            */
/* ===================================================================== */

/* relline: lineE Index1 Index2 Rbr a Incl Rin Rout z limb
 * (src/modelfiles/lmodel_relxill_public.dat:1-11) */
void lmodrelline(const double *energy, int Nflux, const double *parameter, int spectrum,
                 double *flux, double *fluxError, const char *init);

/* relline_lp: lineE h a Incl Rin Rout z limb gamma switch_returnrad
 * (lmodel_relxill_public.dat:23-33); switch_returnrad 1 adds the returning
 * radiation to the lamp-post emissivity (src/Rellp.cpp:483-507) */
void lmodrellinelp(const double *energy, int Nflux, const double *parameter, int spectrum,
                   double *flux, double *fluxError, const char *init);

/* relconv (con model, flux is input and output): Index1 Index2 Rbr a Incl Rin Rout limb
 * (lmodel_relxill_public.dat:13-21) */
void lmodrelconv(const double *energy, int Nflux, const double *parameter, int spectrum,
                 double *flux, double *fluxError, const char *init);

/* relconv_lp: h beta a Incl Rin Rout limb gamma switch_returnrad
 * (lmodel_relxill_public.dat:35-44); switch_returnrad as for relline_lp */
void lmodrelconvlp(const double *energy, int Nflux, const double *parameter, int spectrum,
                   double *flux, double *fluxError, const char *init);

/* relxill: Index1 Index2 Rbr a Incl Rin Rout z gamma logxi Afe Ecut refl_frac
 * (lmodel_relxill_public.dat:46-59; the reference's c_lmodrelxill,
 * LocalModel::relxill_model -> relxill_kernel, src/LocalModel.cpp:53-72,
 * src/Relxill.cpp:289-414): the xillver spectra of the table interpolated in
 * Gamma, A_Fe, logXi, Ecut for every emission angle, weighted by the angle
 * distribution, convolved with the relativistic profile, times |refl_frac|,
 * plus the cut-off power law for refl_frac >= 0. */
void lmodrelxill(const double *energy, int Nflux, const double *parameter, int spectrum,
                 double *flux, double *fluxError, const char *init);

/* relxilllp: h beta a Incl Rin Rout z gamma logxi Afe Ecut refl_frac
 * switch_returnrad switch_reflfrac_boost (lmodel_relxill_public.dat:61-75; the
 * reference's c_lmodrelxilllp -> relxill_kernel with 10 radial zones,
 * src/relutility.c:506-544): per zone the relativistic profile and angle
 * distribution, the xillver spectra for the Ecut the zone sees (source ->
 * disk energy shift, src/IonGradient.cpp:182-209) divided by the change of
 * the incident normalisation (src/Xillspec.cpp:408-435), one convolution per
 * zone; the reflection scaled by |refl_frac| over the lamp post's reflection
 * fraction (by |refl_frac| alone with switch_reflfrac_boost = 1) and the
 * primary boosted by f_inf / 0.5 g^Gamma (src/PrimarySource.cpp:66-125).
 * switch_returnrad 1 (the lmodel.dat default) adds the returning radiation
 * to the emissivity, above spin 0 corrected by each zone's xillver flux and
 * energy-shift factors (src/Relxill.cpp:337-344, src/Xillspec.cpp:500-526);
 * 2 and -1 use the returning part alone. */
void lmodrelxilllp(const double *energy, int Nflux, const double *parameter, int spectrum,
                   double *flux, double *fluxError, const char *init);

/* relxillCp: Incl a Rin Rout Rbr Index1 Index2 z gamma logxi logN Afe kTe
 * refl_frac (lmodel_relxill_public.dat:96-110; the reference's
 * c_lmodrelxilldensnthcomp): relxill with the nthcomp primary
 * (spec_nthcomp, src/Xillspec.cpp:243-253; kT_bb 0.05 keV) and the nthcomp
 * xillver table, interpolated in Gamma, A_Fe, logXi, kTe (and logN when the
 * table has a density axis). */
void lmodrelxilldensnthcomp(const double *energy, int Nflux, const double *parameter, int spectrum,
                            double *flux, double *fluxError, const char *init);

/* relxilllpCp: Incl a Rin Rout h beta gamma logxi logN Afe kTe refl_frac z
 * iongrad_index iongrad_type switch_returnrad switch_reflfrac_boost
 * (lmodel_relxill_public.dat:112-129; the reference's
 * c_lmodrelxilllpdensnthcomp): relxilllp with the nthcomp primary.  kTe is
 * the source's (not shifted to the observer as relxilllp's Ecut,
 * src/Relxill.cpp:197-200); each zone sees kTe times its source -> disk
 * shift, and its normalisation change is that of the nthcomp spectrum.
 * iongrad_type must be 0 (constant ionisation; the gradients return
 * RX_ERR_UNSUPPORTED). */
void lmodrelxilllpdensnthcomp(const double *energy, int Nflux, const double *parameter, int spectrum,
                              double *flux, double *fluxError, const char *init);

/* relxillNS: Index1 Index2 Rbr a Incl Rin Rout z kTbb logxi Afe logN
 * refl_frac (lmodel_relxill_public.dat:140-153; the reference's
 * c_lmodrelxillns): relxill (broken power law, one zone) with the blackbody
 * primary E^2 / (kT^4 (exp(E/kT) - 1)) (spec_blackbody,
 * src/Xillspec.cpp:255-268), normalised as the xillver incident spectrum, and
 * the blackbody xillver table interpolated in kTbb, A_Fe, logXi, logN. */
void lmodrelxillns(const double *energy, int Nflux, const double *parameter, int spectrum,
                   double *flux, double *fluxError, const char *init);

/* The non-relativistic reflection models (LocalModel::xillver_model,
 * src/LocalModel.cpp:104-128): the xillver table (the model's own: the three
 * tables above) interpolated in all five parameters, the inclination too
 * (interp_5d_tab, src/xilltable.c:878-995), times 0.5 cos(Incl)
 * (norm_xillver_spec, src/Xillspec.cpp:536-545), rebinned onto the
 * redshifted grid, times |refl_frac|, plus the normalised primary on that
 * grid for refl_frac >= 0 (add_primary_component, src/Relbase.cpp:294-344).
 *   xillver: gamma Afe Ecut logxi z Incl refl_frac (lmodel_relxill_public.dat:77-84)
 *   xillverCp: gamma Afe kTe logxi logN z Incl refl_frac (:86-94;
 *     the reference's c_lmodxillverdensnthcomp)
 *   xillverNS: kTbb Afe logN logxi z Incl refl_frac (:131-138) */
void lmodxillver(const double *energy, int Nflux, const double *parameter, int spectrum,
                 double *flux, double *fluxError, const char *init);
void lmodxillverdensnthcomp(const double *energy, int Nflux, const double *parameter, int spectrum,
                            double *flux, double *fluxError, const char *init);
void lmodxillverns(const double *energy, int Nflux, const double *parameter, int spectrum,
                   double *flux, double *fluxError, const char *init);

/* Status-returning evaluation of one model (the body of xspec_C_wrapper_eval_model,
 * src/LocalModel.cpp:143-160, and LocalModel::eval_model, src/LocalModel.h:108-128).
 * precision: RX_FP64 or RX_FP32. */
int rx_eval_model(int model, const double *parameter, const double *energy, double *flux,
                  int n_flux, int precision);

int rx_last_status(void);
const char *rx_last_error(void);

/* Allocation-free variant of rx_eval_model.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  One request of relline, relline_lp,
 * relconv or relconv_lp on device data the caller owns: d_energy (n_flux+1
 * edges), d_flux_in (the input spectrum of the convolution models, n_flux;
 * may equal d_flux_out; unused by the line models), d_flux_out (n_flux);
 * the parameter array in host memory.  The kernels' work buffers come from a
 * fixed device array of RX_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; a request that needs more (grid size, table dims, fine grid)
 * is rejected with RX_ERR_ARG.  Nothing is allocated: the library must be
 * initialised (rx_init) with its tables resident (a changed table is
 * uploaded by rx_init or any other entry point, not by this one).  Runs on
 * `stream` (a hipStream_t; NULL: the null stream) and returns once it is
 * done, with the request's status.  Every call on a GPU uses the same
 * array: calls that overlap on one GPU race. */
#define RX_NOALLOC_SCRATCH_BYTES (4u << 20)
int rx_eval_model_noalloc(int model, const double *parameter, const double *d_energy, const double *d_flux_in,
                          double *d_flux_out, int n_flux, int precision, void *stream);

/* Processed parameters of one XSPEC parameter array (get_rel_params with
 * check_parameter_bounds, src/ModelDefinition.cpp:168-334): out12 receives
 * a, incl[rad], Index1, Index2, Rbr, Rin, Rout, lineE, z, h, gamma, beta. */
int rx_model_params(int model, const double *parameter, double *out12);

/* ===================================================================== */
/* library lifecycle and tables                                          */
/* ===================================================================== */

/* Selects the device and uploads the tables (lazily done by the first call
 * otherwise).  Tables: with RELXILL_TABLE_PATH set, the reference's FITS
 * files rel_table_v0.5a.fits and rel_lp_table_v0.5b.fits (src/common.h:73,80)
 * in that directory, else rel_table.rxtb and rel_lp_table.rxtb (own binary
 * format) there; a table that is missing or cannot be read there makes every
 * request that needs it fail with RX_ERR_TABLE and the reason in
 * rx_last_error(), as the reference fails when it cannot open a table
 * (src/reltable.c:255-268).  With RELXILL_TABLE_PATH unset: seeded
 * synthetic tables at the dims the benchmarks use (25x22x100x20 relline,
 * 20x100x100 lamp post). */
int rx_init(int device);
void rx_shutdown(void);

/* Replace the resident tables with seeded synthetic ones of the given dims
 * (host-only work plus an upload when a device is initialised). */
int rx_tables_synthetic(uint64_t seed, int n_a, int n_mu0, int n_r, int n_g);
int rx_lptable_synthetic(uint64_t seed, int n_a, int n_h, int n_r);
/* Load/save tables in the library's binary format (RXTB). */
int rx_tables_load(const char *path);
int rx_lptable_load(const char *path);
int rx_tables_save(const char *path);
int rx_lptable_save(const char *path);
/* Load the reference's FITS tables: the relline table in the layout read by
 * read_relline_table (src/reltable.c:121-304: extensions "a", "mu0", node
 * (i, j) at HDU i*n_mu0 + j + 4 with columns r, gmin, gmax, trff1, trff2,
 * cosne1, cosne2) and the lamp-post table in the layout read by read_lp_table
 * (src/reltable.c:313-448: extension "I_h", one row per spin with a, hgrid,
 * r, h<k>, del<k>, del_inc<k>; angles stored as absolute values, :370-377).
 * Dimensions come from the file.  cfitsio is loaded at run time from
 * RELXILL_CFITSIO_LIB, $HEADAS/lib, the loader path or $CONDA_PREFIX/lib;
 * without it, or on a malformed file, RX_ERR_TABLE and rx_last_error(). */
int rx_tables_load_fits(const char *path);
int rx_lptable_load_fits(const char *path);

/* xillver reflection table (five or six parameters, the last the
 * inclination; the reference's xillver-a-Ec5.fits, src/common.h:163; a sixth
 * parameter is the first and is interpolated outside the other five, as in
 * interp_6d_tab_incl, src/xilltable.c:997-1045).  Loaded from
 * RELXILL_TABLE_PATH/xillver-a-Ec5.fits on the first relxill request when
 * present, else a synthetic table (13 Gamma x 5 A_Fe x 8 logXi x 6 Ecut x 10
 * inclinations, 2999 log bins over 0.07-1000 keV).  The FITS loader reads
 * the layout of src/xilltable.c:169-277, :513-566 (extensions PARAMETERS with
 * NAME/NUMBVALS/VALUE in columns 1/9/10, ENERGIES, SPECTRA with the spectrum
 * in column 2); spectra are divided by 10^logXi when uploaded
 * (renorm_xill_spec, src/xilltable.c:478-511). */
int rx_xilltable_synthetic(uint64_t seed, int n_gam, int n_afe, int n_lxi, int n_ect, int n_incl, int n_ener);
int rx_xilltable_load_fits(const char *path);
typedef struct {
  int n_par;
  int pid[6];   /* parameter ids: 0 Gamma, 1 A_Fe, 2 logXi, 3 Ecut (kTe), 4 Dens, 5 kTbb, 7 Incl */
  int n_val[6];
  const float *val[6];
  int n_ener;
  const float *elo, *ehi;
  const float *spec;  /* rows as stored in the file (before the 10^logXi division) */
} rx_xilltable_view;
int rx_xilltable_get_view(rx_xilltable_view *v);

/* The nthcomp models' xillver table (kTe in the place of Ecut: the same
 * parameter id 3; the reference's xillverCp_v3.4.fits, src/common.h:166,
 * get_init_xillver_table, src/xilltable.c:733-740), held beside the one
 * above.  Loaded from RELXILL_TABLE_PATH/xillverCp_v3.4.fits on the first
 * relxillCp / relxilllpCp request when present, else a synthetic table as
 * above with its fourth axis kTe 1-400 keV.  Five- and six-parameter tables
 * load; kTe may not be the first of six.
 * rx_xillcptable_synthetic_dens: a seeded six-parameter table, Gamma, A_Fe,
 * logXi, kTe, Dens (15-19), Incl; no counterpart in the reference (a test
 * table for the six-parameter path). */
int rx_xillcptable_synthetic(uint64_t seed, int n_gam, int n_afe, int n_lxi, int n_kte, int n_incl, int n_ener);
int rx_xillcptable_synthetic_dens(uint64_t seed, int n_dens, int n_gam, int n_afe, int n_lxi, int n_kte, int n_incl,
                                  int n_ener);
int rx_xillcptable_load_fits(const char *path);
int rx_xillcptable_get_view(rx_xilltable_view *v);

/* The blackbody model's xillver table (parameters kTbb, A_Fe, logXi, Dens,
 * Incl by name; the reference's xillverNS-2.fits, src/common.h:167,
 * get_init_xillver_table, src/xilltable.c:742-748), held beside the others.
 * Loaded from RELXILL_TABLE_PATH/xillverNS-2.fits on the first relxillNS
 * request when present, else a synthetic table (12 kTbb 0.5-10 keV x 5 A_Fe
 * x 8 logXi 1-4.7 x 5 Dens 15-19 x 10 inclinations, 2999 log bins). Rows are
 * divided by 10^logXi and 10^(Dens - 15) of their node when uploaded.
 * Five-parameter tables only. */
int rx_xillnstable_synthetic(uint64_t seed, int n_ktbb, int n_afe, int n_lxi, int n_dens, int n_incl, int n_ener);
int rx_xillnstable_load_fits(const char *path);
int rx_xillnstable_get_view(rx_xilltable_view *v);

/* Returning-radiation table of the lamp-post models (switch_returnrad = 1;
 * the reference's table_returnRad_v20220301.fits, src/common.h:126-128):
 * loaded from RELXILL_TABLE_PATH on the first request that needs it, else a
 * synthetic table (21 spins x 50 radial zones x 20 energy-shift bins).  The
 * FITS loader reads the layout of src/Relreturn_Table.cpp:126-340 (extension
 * SPIN, then FRAC01, FRAC02, ... with rlo, rhi, tf_r, gmin, gmax, frac_g). */
int rx_rradtable_synthetic(uint64_t seed, int n_spin, int n_rad, int n_g);
int rx_rradtable_load_fits(const char *path);
typedef struct {
  int n_spin, n_rad, n_g;
  const double *spin, *rlo, *rhi, *tf_r, *gmin, *gmax, *frac_g;
} rx_rradtable_view;
int rx_rradtable_get_view(rx_rradtable_view *v);

/* Read-only host view of the resident tables (for checking). */
typedef struct {
  int n_a, n_mu0, n_r, n_g;
  const float *a, *mu0, *r, *gmin, *gmax, *trff1, *trff2, *cosne1, *cosne2;
} rx_reltable_view;
typedef struct {
  int n_a, n_h, n_r;
  const float *a, *h, *rad, *intens, *del, *del_inc;
} rx_lptable_view;
int rx_tables_view(rx_reltable_view *rel, rx_lptable_view *lp);

/* Convolution grid of relconv (reference: 4096 log bins 0.00035-2000 keV,
 * src/Xillspec.h:36-38, src/Xillspec.cpp:44-54).  n must be a power of two
 * (<= 8192).  Fine radial grid size (reference N_FRAD = 1000). */
int rx_set_conv_grid(int n, double emin, double emax);
int rx_set_nr_fine(int nr_fine);
/* FP32 path: relative guard band around the Romberg stop threshold (default
 * 5e-4); stop tests inside it are redone in FP64.  Widening it exercises the
 * FP64 re-check path (diagnostics). */
int rx_set_fp32_guard(double rel);

/* ===================================================================== */
/* batched API: inputs resident in HBM, many parameter sets and uneven    */
/* energy grids per launch (project-defined; no reference counterpart)    */
/* ===================================================================== */

typedef struct rx_batch rx_batch;

typedef struct {
  int model;               /* RX_MODEL_* */
  int n_flux;              /* bins of the caller grid */
  const double *energy;    /* n_flux+1 edges (host memory, copied at create) */
  const double *parameter; /* XSPEC parameter array of the model (copied) */
  const double *flux_in;   /* conv models: input spectrum n_flux (copied) */
  int refl_set;            /* RX_MODEL_RELXILL_CONV: reflection set id */
} rx_request;

/* Builds a device-resident batch: copies grids, parameters and inputs to HBM,
 * validates parameters (per-request status), sizes the work list.  NULL on failure. */
rx_batch *rx_batch_create(const rx_request *req, int n_req);
/* Runs the whole path on the device (stream-ordered, asynchronous). */
int rx_batch_run(rx_batch *b, int precision);
int rx_batch_sync(rx_batch *b);
/* Copies request i's output spectrum (n_flux values) to host memory and
 * returns its status. */
int rx_batch_output(rx_batch *b, int i, double *flux);
int rx_batch_status(rx_batch *b, int i);
/* emission-angle distribution of a relxill-style or relxill request: the
 * first n values of [zone][n_incl] (one zone except relxilllp's 10) */
int rx_batch_dist(rx_batch *b, int i, double *dist, int n);
/* radial zones of a relxill / relxilllp request after a run, 8 values per
 * zone: mean radius, emission angle of the zone (relxilllp), energy shift
 * source -> zone, xillver normalisation ratio nf(Ecut_zone)/nf(Ecut_source),
 * lower Ecut table index of the zone and its interpolation factor, lower
 * logXi table index and factor (the zones of relxill_kernel,
 * src/relutility.c:506-544; Ecut shift, src/Relxill.cpp:357-388; the logXi
 * of an ionisation gradient, src/IonGradient.cpp:340-395) */
int rx_batch_zones(rx_batch *b, int i, double *out, int n_zones);
/* lamp-post reflection fractions of request i: refl_frac, f_bh, f_ad, f_inf, f_inf_rest */
int rx_batch_reflfrac(rx_batch *b, int i, double *out5);
/* Integration work shape: bins per work item (segment of one chunk of 64
 * fine radii; 0: automatic / autotuned) and persistent wavefronts per CU
 * (0: automatic / autotuned). */
int rx_batch_set_shape(rx_batch *b, int seg_bins, int waves_per_cu);
void rx_batch_destroy(rx_batch *b);

/* Resident reflection set for relxill-style requests: n_incl spectra of
 * n_refl bins (photons/bin) on refl_energy (n_refl+1 edges).  Returns the set
 * id (>= 0) or a negative status. */
int rx_reflset_create(const double *refl_energy, int n_refl, const double *refl, int n_incl);
/* Frees every resident reflection set (after all batches using them are done);
 * set ids start again from 0.  Needed before changing the conv grid size. */
int rx_reflsets_clear(void);

/* Kernel timing of the last rx_batch_run with profiling enabled: per kernel
 * (0 k_setup (with k_reset), 1 k_integrate, 2 k_fixup, 3 k_finalize, 4 k_conv,
 * 5 k_xill, 6 k_primary) the duration in ms measured with HIP events on the
 * batch's stream. */
#define RX_N_TIMED 7
int rx_profile_enable(int on);
/* Streams `bytes` through a plain device copy kernel `reps` times (after one
 * warm-up) and returns the average kernel time; used to calibrate HBM byte
 * counters and as a bandwidth reference. */
int rx_copy_test(int64_t bytes, int reps, double *ms);
/* Waits for all work of the library on its device. */
int rx_device_synchronize(void);

/* The nthcomp primary continuum (spec_nthcomp, src/Xillspec.cpp:243-253, with
 * c_donthcomp, src/donthcomp.c:683-793): n_spec spectra on one grid of
 * n_flux bins (energy: n_flux + 1 ascending positive edges, keV), spectrum s
 * with photon index gam[s], electron temperature kte[s] (keV, in the frame of
 * the source) and energy shift source -> observer ener_shift[s] (redshift
 * 1 / ener_shift - 1), seed a 0.05 keV disc blackbody.  flux receives n_spec
 * rows of n_flux photons per bin, normalised to E F_E = 1 at 1 keV in the
 * source frame.  One Kompaneets solve per spectrum on the device. */
int rx_nthcomp(const double *energy, int n_flux, int n_spec, const double *kte, const double *gam,
               const double *ener_shift, double *flux);
int rx_batch_kernel_ms(rx_batch *b, double *ms);
/* Work counters of the last completed rx_batch_run (after rx_batch_sync):
 * out[0] work items integrated, out[1] FP32 bins re-checked in FP64,
 * out[2] 1 if the re-check queue overflowed (the batch was then redone in
 * FP64).  No counterpart in the reference (diagnostics of the batched path). */
int rx_batch_counters(rx_batch *b, int64_t *out);

/* Launch graphs (default on): a batch run whose kernel arguments and grid
 * shapes equal the previous run's is captured into a hipGraph and replayed
 * while they stay equal (the XSPEC entry points reuse one batch, so repeated
 * calls on one grid replay a graph).  rx_set_graphs(0) launches every kernel
 * directly.  rx_batch_graph_replays: runs of the batch served by a graph.
 * No counterpart in the reference (launch mechanics of the device path). */
int rx_set_graphs(int on);
int64_t rx_batch_graph_replays(rx_batch *b);

/* Autotuner: times candidate work shapes (segment bins x waves per CU) for
 * this batch on the device and keeps the fastest for batches of the same
 * shape class.  Returns 256 * seg_bins + waves_per_cu (>0) or -status. */
int rx_batch_autotune(rx_batch *b, int precision);

/* ===================================================================== */
/* serving loop: continuous batching with HBM admission                  */
/* ===================================================================== */

typedef struct rx_server rx_server;
/* max_inflight: requests on the device at once; hbm_budget_bytes: admission
 * cap for per-request device memory; n_streams: concurrent batches. */
rx_server *rx_server_start(int max_inflight, int64_t hbm_budget_bytes, int n_streams,
                           int precision);
/* Submits a request (copied); returns a ticket >= 0 or a negative status. */
int64_t rx_server_submit(rx_server *s, const rx_request *req);
/* Load generator for measuring the server itself: submits req[i] at
 * arrival_s[i] seconds after the call (ascending), paced by a native loop so
 * that the offered rate does not depend on the caller's language; tickets[i]
 * receives the ticket (or a negative status).  Returns once the last request
 * is submitted; poll or drain for the results. */
int rx_server_replay(rx_server *s, const rx_request *req, int n, const double *arrival_s, int64_t *tickets);
/* Non-blocking poll: 1 done (flux copied, *status set, *latency_ms set),
 * 0 pending, <0 unknown ticket. */
int rx_server_poll(rx_server *s, int64_t ticket, double *flux, int *status, double *latency_ms);
/* Blocks until every submitted request completed. */
int rx_server_drain(rx_server *s);
/* counters: completed, batches launched, max concurrent batches */
int rx_server_stats(rx_server *s, int64_t *out3);
void rx_server_stop(rx_server *s);

#ifdef __cplusplus
}
#endif
#endif
