/*
 * relxill_oracle.h -- FP64 CPU oracle for the relativistic smearing path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of relxill 2.4's
 * relline / relline_lp / relconv algorithm (the reference is C/C++; every
 * function in relxill_oracle.c cites the reference file:line it follows).
 * It is linked only by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, always as the checker or the timed CPU baseline, never as
 * part of the product path.
 *
 * Pinning: tests/test_oracle_kats.py checks the functions against the
 * reference's known-answer tests and against closed forms, including the
 * ISCO / 1000 rg snaps of the table interpolation and the 75 deg lin/log
 * switch of the lamp-post radius interpolation.
 *
 * Conventions follow the reference:
 *   - energy grids have n+1 ascending edges, fluxes n values (photons/bin);
 *   - relline tables are float32, node-major [ia][imu0][ir](.. [ig]);
 *   - radii on the fine grid are descending (rout -> rin).
 */
#ifndef RELXILL_ORACLE_H_
#define RELXILL_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* model types: same numeric ids as the reference (src/common.h:85-95) */
#define ORC_MOD_RELLINE 1
#define ORC_MOD_RELLINELP 2
#define ORC_MOD_RELCONV 11
#define ORC_MOD_RELCONVLP 12
#define ORC_MOD_RELXILL (-1)
#define ORC_MOD_RELXILLLP (-2)
/* project id: relxill-style convolution of a given angle-resolved reflection
 * set, with relconv's 8 parameters and relxill's profile normalisation */
#define ORC_MOD_RELXILL_CONV (-1001)
/* relxillCp / relxilllpCp: relxill / relxilllp with the nthcomp primary
 * (parameters of lmodel_relxill_public.dat:96-129) */
#define ORC_MOD_RELXILLCP (-1010)
#define ORC_MOD_RELXILLLPCP (-1011)
/* relxillNS: relxill with the blackbody primary (lmodel_relxill_public.dat:140-153) */
#define ORC_MOD_RELXILLNS (-1012)
/* the non-relativistic xillver models (MOD_TYPE_XILLVER, _NTHCOMP, NS,
 * src/common.h:91-111; lmodel_relxill_public.dat:77-138) */
#define ORC_MOD_XILLVER 0
#define ORC_MOD_XILLVERCP 100
#define ORC_MOD_XILLVERNS (-101)

/* emissivity types (src/common.h:50-55) */
#define ORC_EMIS_BKN 1
#define ORC_EMIS_LP 2

/* relline transfer-function table (src/reltable.h:24-48), dims at run time */
typedef struct {
  int n_a, n_mu0, n_r, n_g;
  const float *a;      /* [n_a] ascending spin */
  const float *mu0;    /* [n_mu0] ascending cos(incl) */
  const float *r;      /* [n_a][n_mu0][n_r] descending */
  const float *gmin;   /* [n_a][n_mu0][n_r] */
  const float *gmax;   /* [n_a][n_mu0][n_r] */
  const float *trff1;  /* [n_a][n_mu0][n_r][n_g] */
  const float *trff2;
  const float *cosne1;
  const float *cosne2;
} orc_reltable;

/* returning-radiation table (src/Relreturn_Table.h): per spin n_rad zones
 * [rlo, rhi] and per (incident ro, emitting re) pair tf_r, gmin, gmax and
 * frac_g[n_g]; flat, spin-major */
typedef struct orc_rradtable {
  int n_spin, n_rad, n_g;
  const double *spin;            /* [n_spin] ascending */
  const double *rlo, *rhi;       /* [n_spin][n_rad] */
  const double *tf_r, *gmin, *gmax;  /* [n_spin][n_rad][n_rad] */
  const double *frac_g;          /* [n_spin][n_rad][n_rad][n_g] */
} orc_rradtable;

/* lamp-post table (src/reltable.h:51-69): the h grid is stored per spin */
typedef struct {
  int n_a, n_h, n_r;
  const float *a;        /* [n_a] ascending */
  const float *h;        /* [n_a][n_h] ascending */
  const float *rad;      /* [n_a][n_r] ascending */
  const float *intens;   /* [n_a][n_h][n_r] */
  const float *del;      /* [n_a][n_h][n_r], already |.| as at load time */
  const float *del_inc;  /* [n_a][n_h][n_r] */
  const orc_rradtable *rrad;  /* returning radiation (may be NULL: not available) */
} orc_lptable;

/* relativistic parameters (src/common.h:191-214, subset on this path) */
typedef struct {
  int model_type;
  int emis_type;
  double a, incl; /* incl in rad */
  double emis1, emis2, rbr;
  double rin, rout;
  double lineE, z;
  double height, htop, gamma, beta;
  int limb;
  int num_zones;
  int return_rad;
  int do_renorm;
  /* relxilllp returning radiation: the xillver flux and energy-shift
   * correction factors of the rr_corr_n radial zones with edges
   * rr_corr_rgrid[rr_corr_n+1] (rrad_corr_factors, src/common.h:213);
   * rr_corr_n = 0: no correction */
  const double *rr_corr_flux, *rr_corr_gshift, *rr_corr_rgrid;
  int rr_corr_n;
  /* relxilllpCp's ionisation gradient (src/IonGradient.cpp:340-395): type
   * (0 constant, 1 power law), index and the input logXi it starts from */
  int ion_grad_type;
  double ion_grad_index, lxi0;
} orc_relparam;

/* reflection fractions (src/common.h:279-285) */
typedef struct {
  double refl_frac, f_bh, f_ad, f_inf, f_inf_rest;
} orc_reflfrac;

/* per-call work counters (optional) */
typedef struct {
  int64_t relb_func_calls;
  int64_t romberg_calls;
  int64_t romberg_levels;
  int64_t edge_calls;
  int64_t midpoint_calls;
  int64_t bin_integrals;
} orc_stats;

/* ---- physics helpers (src/Relphysics.cpp) ---- */
double orc_kerr_rms(double a);
double orc_kerr_rplus(double a);
double orc_ut_disk(double r, double a);
double orc_gi_potential_lp(double r, double a, double h, double bet, double del);
double orc_lp_emissivity_newton(double h, double r);

/* ---- grids and utilities (src/relutility.c) ---- */
void orc_get_log_grid(double *ener, int n_edges, double emin, double emax);
void orc_get_fine_radial_grid(double rin, double rout, double *re, int nr);
int orc_binary_search(const double *arr, int n, double val);
int orc_binary_search_float(const float *arr, int n, float val);
int orc_inv_binary_search(const double *arr, int n, double val);
double orc_interp_lin_1d(double ifac, double rlo, double rhi);
double orc_interp_lin_2d(double i1, double i2, double r11, double r12, double r21, double r22);
double orc_interp_lin_2d_float(double i1, double i2, float r11, float r12, float r21, float r22);
void orc_rebin_spectrum(const double *ener, double *flu, int nbins,
                        const double *ener0, const double *flu0, int nbins0);
void orc_rebin_mean_flux(const double *xn, double *yn, int nn,
                         const double *x0, const double *y0, int n0);
int orc_get_rzone_grid(double rmin, double rmax, int nzones, double h, double *rgrid);

/* ---- parameters (src/ModelDefinition.cpp:153-334, src/LocalModel.cpp) ---- */
/* Fills *p from the XSPEC parameter array of model `model_type`; returns 0 on
 * success, nonzero when the reference would reject the parameters. */
int orc_get_rel_params(int model_type, const double *parameter, int phys_norm_env,
                       orc_relparam *p);
int orc_n_params(int model_type);

/* ---- system parameters (src/Relprofile.cpp:141-303) ---- */
typedef struct {
  int nr, ng;
  double *re, *gmin, *gmax;   /* [nr] */
  double *gstar, *d_gstar;    /* [ng] */
  double *trff, *cosne;       /* [nr][ng][2] */
  double *emis, *del_emit, *del_inc; /* [nr] */
  int limb_law;
  orc_reflfrac refl;
} orc_syspar;

orc_syspar *orc_syspar_new(int nr, int ng);
void orc_syspar_free(orc_syspar *s);
int orc_interpol_reltable(const orc_reltable *tab, double a, double incl, double rin,
                          double rout, orc_syspar *out);
int orc_calc_emis_profile(const orc_lptable *lp, const orc_relparam *p, orc_syspar *s);
/* returning-radiation emissivity on the fine grid re (descending) for the
 * direct emissivity emis0 (get_rrad_emis_corona, src/Relreturn_Corona.cpp:291-321,
 * without the relxill correction factors) */
int orc_rrad_emis(const orc_rradtable *rt, const orc_relparam *p, const double *re, const double *emis0, int nr,
                  double *emis_ret);
/* everything get_system_parameters does (src/Relprofile.cpp:310-358) */
orc_syspar *orc_get_system_parameters(const orc_reltable *tab, const orc_lptable *lp,
                                      const orc_relparam *p, int nr_fine, int *status);

/* ---- relline profile (src/Relprofile.cpp:835-958, :749-796) ---- */
/* flux: [nzones][n_ener]; dist (optional, may be NULL): [nzones][n_incl] */
int orc_calc_relline_profile(const orc_syspar *s, const double *ener, int n_ener,
                             const double *rgrid, int nzones, double *flux,
                             double *dist, int n_incl, orc_stats *st);
void orc_renorm_relline_profile(const double *ener, int n_ener, int nzones, double *flux,
                                double *dist, int n_incl, const orc_relparam *p);
/* relbase (src/Relbase.cpp:524-538): syspar + zones + profile + renorm */
int orc_relbase(const orc_reltable *tab, const orc_lptable *lp, const orc_relparam *p,
                const double *ener, int n_ener, int nr_fine, double *flux,
                double *dist, int n_incl, orc_stats *st);

/* ---- FFT convolution (src/Relbase.cpp:93-246) ---- */
void orc_fft(double *re, double *im, int n, int inverse);
void orc_fft_conv_spectrum(const double *ener, const double *fxill, const double *frel,
                           double *fout, int n);
double orc_fft_norm_factor(const double *ener, const double *fxill, const double *frel,
                           const double *fout, int n);
void orc_conv_normalized(const double *ener, const double *fxill, const double *frel,
                         double *fout, int n);
void orc_zero_outside(const double *ener, double *spec, int n, double emin, double emax);
/* relconv_kernel (src/Relbase.cpp:257-289) on a given convolution grid */
int orc_relconv_kernel(const orc_reltable *tab, const orc_lptable *lp, const orc_relparam *p,
                       const double *conv_ener, int n_conv, const double *ener_inp,
                       double *spec_inp, int n_inp, int nr_fine, orc_stats *st);

/* relxill-style angle-resolved convolution (src/Relxill.cpp:432-482 with
 * src/Xillspec.cpp:557-573), one zone: refl[n_incl][n_refl] on refl_ener. */
int orc_relxill_conv(const orc_reltable *tab, const orc_relparam *p, const double *conv_ener,
                     int n_conv, const double *refl_ener, const double *refl, int n_refl,
                     int n_incl, const double *out_ener, double *out, int n_out, int nr_fine,
                     double *dist_out, orc_stats *st);

/* ---- xillver table and the relxill model ---- */
/* xillver parameter ids (PARAM_*, src/common.h:141-149) */
#define ORC_XP_GAM 0
#define ORC_XP_AFE 1
#define ORC_XP_LXI 2
#define ORC_XP_ECT 3
#define ORC_XP_DNS 4
#define ORC_XP_KTB 5
#define ORC_XP_INC 7

/* a five- or six-parameter xillver table (xillTable, src/xilltable.h), the
 * last parameter the inclination; spec holds the rows as stored in the file
 * (row r at spec[r * n_ener], r = ((((i0 n1 + i1) n2 + i2) n3 + i3) n4 + i4)
 * for five, one more index in front for six, src/xilltable.c:326-337), before
 * the division by 10^logXi */
typedef struct {
  int n_par;
  int pid[6];
  int n_val[6];
  const float *val[6];
  int n_ener;
  const float *elo, *ehi;
  const float *spec;
} orc_xilltable;

/* xillver parameters of one evaluation (xillTableParam, src/xilltable.h) */
typedef struct {
  double gam, afe, lxi, ect, dens, incl;  /* incl in degrees */
  double refl_frac;
  int boost;  /* switch_reflfrac_boost: refl_frac relative to the lamp post's */
  int prim;   /* primary continuum: 0 cut-off power law (ect = Ecut), 1 nthcomp (ect = kTe),
                2 blackbody (ktbb) */
  double ktbb; /* blackbody temperature (keV) */
} orc_xillparam;

/* get_xilltab_indices_for_paramvals (src/xilltable.c:297-324) */
void orc_xill_indices(const orc_xilltable *t, const orc_xillparam *x, int *ind);
/* the interpolation factors of interp_xill_table (src/xilltable.c:1090-1125)
 * with resetInpvalsToBoundaries (:1069-1088) and ensure_ecut_within_boundarys
 * (:1054-1067) */
void orc_xill_factors(const orc_xilltable *t, const orc_xillparam *x, const int *ind, double *fac);
/* renorm_xill_spec (src/xilltable.c:478-486) */
void orc_xill_renorm(float *spec, int n, double lxi, double dens);
/* get_xillver_spectra_table for a relxill model (src/Xillspec.cpp:58-88):
 * ener[n_ener+1] (elo, then the last ehi), flu[n_incl][n_ener] */
void orc_xill_spectra(const orc_xilltable *t, const orc_xillparam *x, double *ener, double *flu);
/* get_xillver_fluxcorrection_factors (src/Xillspec.cpp:500-526) of the
 * xillver spectra at the given parameters: *corr_flux (reflected over
 * incident energy flux in 0.1-1000 keV) and *corr_gshift (photon-flux boost
 * ratio to a power law at g = 1.5) */
void orc_xill_corr_factors(const orc_xilltable *t, const orc_xillparam *x, double *corr_flux, double *corr_gshift);
/* interp_5d_tab for the xillver models (src/xilltable.c:878-995): one
 * spectrum, the inclination (degrees) interpolated and clamped too */
void orc_xill_spectrum_incl(const orc_xilltable *t, const orc_xillparam *x, double *ener, double *flu);
/* xillver, xillverCp, xillverNS on a caller grid (LocalModel::xillver_model,
 * src/LocalModel.cpp:104-128); 0 or a status */
int orc_xillver_model(int model_type, const orc_xilltable *xt, const double *par, const double *energy,
                      double *flux, int n_flux);
/* calc_xillver_angdep (src/Xillspec.cpp:557-573) */
void orc_xill_angdep(double *out, const double *flu, const double *dist, int n_incl, int n_ener);
/* spec_cutoffpl (src/Xillspec.cpp:215-241) */
void orc_spec_cutoffpl(double *flux, const double *ener, int n, double ect, double gam, double shift);
/* calcNormWrtXillverTableSpec (src/Xillspec.cpp:179-205); *status = 1 when
 * the grid does not cover 0.1-1000 keV */
double orc_norm_wrt_xillver(const double *flux, const double *ener, int n, int *status);
/* 1 / calcNormWrtXillverTableSpec of the cut-off power law on the coarse
 * 500-bin grid (PrimarySource::calc_normalization_factor_source,
 * src/PrimarySource.h:279-292) */
double orc_primary_norm_factor(double gam, double ect);

/* ---- nthcomp primary continuum (nthcomp_oracle.c; src/donthcomp.c) ---- */
/* diskbb photon shape per unit Tin^2 at E / Tin (f_mcdint__) */
double orc_mcd_shape(double et);
/* Kompaneets solve for a disc seed: x[0..jmax] (m_e c^2), spt[0..jmax-1]
 * (E F_E); arrays of at least 901 / 900; returns jmax */
int orc_nthcomp_solve(double tempbb, double theta, double gamma, double *x, double *spt);
/* photons per bin of ear[0..ne] at redshift z, E F_E(1 keV, source) = 1 */
void orc_nthcomp_photar(const double *x, const double *spt, int nth, const double *ear, int ne, double z,
                        double *photar);
/* calc_primary_spectrum (src/Xillspec.cpp:278-300) for the primary of x with
 * Ecut / kTe = ect and the energy shift `shift` */
void orc_prim_spectrum(double *flux, const double *ener, int n, const orc_xillparam *x, double ect, double shift);
/* 1 / calcNormWrtXillverTableSpec of that primary (unshifted) on the 500-bin
 * 0.1-1000 keV grid (src/Xillspec.cpp:179-205, :302-315) */
double orc_prim_norm_factor(const orc_xillparam *x, double ect);
/* spec_blackbody (src/Xillspec.cpp:255-268) */
void orc_spec_blackbody(double *flux, const double *ener, int n, double ktbb);
/* spec_nthcomp (src/Xillspec.cpp:243-253) */
void orc_spec_nthcomp(double *flux, const double *ener, int n, double kte, double gam, double ener_shift);
/* relxill parameters: Index1 Index2 Rbr a Incl Rin Rout z gamma logxi Afe
 * Ecut refl_frac (src/modelfiles/lmodel_relxill_public.dat:46-59); relxilllp:
 * h beta a Incl Rin Rout z gamma logxi Afe Ecut refl_frac switch_returnrad
 * switch_reflfrac_boost (:61-75) */
int orc_get_relxill_params(int model_type, const double *parameter, int phys_norm_env, orc_relparam *p,
                           orc_xillparam *x);
/* energy shift primary source -> observer (src/Relphysics.cpp:239-254) */
double orc_energy_shift_source_obs(const orc_relparam *p);
/* the zones' logXi (IonGradient::calculate_gradient, src/IonGradient.cpp:
 * 340-395): lxi (the request's, kept within 0-4.7) in every zone, or for the
 * power-law gradient log(exp(lxi0) (rmean / rmean[0])^-index) kept within
 * 0-4.7 (calc_ion_grad_pl, :174-182; :36-44, :388) */
void orc_zone_lxi(const orc_relparam *p, double lxi, const double *rmean, int nz, double *out);
/* relxill_kernel (src/Relxill.cpp:289-414 with src/PrimarySource.cpp:66-125)
 * for relxill (broken power law, one zone) and relxilllp (lamp post, 10
 * zones with the source -> disk energy shift of Ecut and the normalisation
 * change, lamp-post reflection fraction and primary boost); constant
 * ionisation.  dist_out (optional): [num_zones][n_incl]. */
int orc_relxill_model(int model_type, const orc_reltable *tab, const orc_lptable *lp,
                      const orc_xilltable *xt, const double *parameter, const double *energy, double *flux,
                      int n_flux, const double *conv_ener, int n_conv, int phys_norm_env, double *dist_out,
                      orc_stats *st);
/* the zones of a relxill model: per zone {r_lo, r_hi, del_emit at the centre,
 * source -> disk energy shift, Ecut seen by the zone, normalisation change,
 * logXi of the zone} */
int orc_relxill_zones(int model_type, const orc_reltable *tab, const orc_lptable *lp, const double *parameter,
                      double *out);
int orc_relxill(const orc_reltable *tab, const orc_xilltable *xt, const double *parameter,
                const double *energy, double *flux, int n_flux, const double *conv_ener,
                int n_conv, int phys_norm_env, double *dist_out, orc_stats *st);

/* ---- model level: what lmod<name> computes (src/LocalModel.cpp:32-98) ---- */
/* energy: n_flux+1 edges; flux: n_flux (input spectrum for conv models).
 * conv_ener/n_conv: the convolution grid for conv models (NULL -> reference grid). */
int orc_eval_model(int model_type, const orc_reltable *tab, const orc_lptable *lp,
                   const double *parameter, const double *energy, double *flux, int n_flux,
                   const double *conv_ener, int n_conv, int phys_norm_env, orc_stats *st);

/* multi-threaded batch of independent model evaluations (CPU baseline) */
typedef struct {
  int model_type;
  const double *parameter;
  const double *energy;
  double *flux;
  int n_flux;
  int status;
} orc_job;
int orc_eval_batch_mt(const orc_reltable *tab, const orc_lptable *lp, orc_job *jobs, int n_jobs,
                      const double *conv_ener, int n_conv, int n_threads);
/* corrected_gshift_fluxboost_factor (src/Relreturn_Corona.cpp:39-83): the flux
 * boost of the xillver reflection under an energy shift g, from its boost
 * ratio c to a power law at g = 1.5 */
double orc_rrad_gshift_boost(double c, double g, double gamma);
/* the same, summing the jobs' work counters into *total (when not NULL) */
int orc_eval_batch_mt_stats(const orc_reltable *tab, const orc_lptable *lp, orc_job *jobs, int n_jobs,
                            const double *conv_ener, int n_conv, int n_threads, orc_stats *total);

#ifdef __cplusplus
}
#endif
#endif
