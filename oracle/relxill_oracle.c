/*
 * relxill_oracle.c -- FP64 CPU oracle of relxill's relativistic smearing path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, as the checker / the timed CPU baseline.  It is
 * never called by the product library.
 *
 * This file restates, in plain C with run-time table dimensions, no global
 * caches and full reentrancy, the algorithm of relxill 2.4 (C/C++).  Each
 * function names the reference lines it follows.  The reference's quirks are
 * kept on purpose (float interpolation factors, swapped lerp weights in
 * relb_func, the Romberg stop rule, the red-wing midpoint switch, ISCO and
 * 1000 rg snaps, rebin edge semantics).  Equivalent restatements:
 *   - relb_func's g-index cache (src/Relprofile.cpp:497-499) always yields the
 *     result of binary_search, so the search is done directly;
 *   - the Romberg endpoint cache is never hit (cache_bin_ener is only ever set
 *     to -1, src/Relprofile.cpp:717,739), so endpoints are always evaluated;
 *   - FFTW's unnormalised r2c/c2r pair (src/Relbase.cpp:153-182) is replaced by
 *     an iterative radix-2 complex FFT with the same sign and scale conventions,
 *     the c2r input rebuilt from the n/2+1 Hermitian half as FFTW reads it.
 * Compile with -ffp-contract=off so that every double operation rounds as
 * written.
 */
#include "relxill_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static const double GFAC_H = 5e-3;            /* src/Relprofile.cpp:36 */
static const double RELTABLE_MAX_R = 1000.0;  /* src/common.h:72 */
static const double EMIN_BAND = 0.01;         /* src/Xillspec.h:30 (EMIN_XILLVER) */
static const double EMAX_BAND = 1000.0;       /* src/Xillspec.h:28,31 */
static const double RELCONV_EMIN = 0.01;      /* src/Relbase.h:47 */
static const double RELCONV_EMAX = 1000.0;    /* src/Relbase.h:48 */

/* ------------------------------------------------------------------------ */
/* physics helpers                                                           */
/* ------------------------------------------------------------------------ */

/* ISCO radius of a Kerr black hole (Bardeen, Press & Teukolsky 1972);
 * src/Relphysics.cpp:139-150 */
double orc_kerr_rms(double a) {
  double sgn = (a < 0) ? -1.0 : 1.0;
  double z1 = 1.0 + pow(1.0 - a * a, 1.0 / 3.0) * (pow(1.0 + a, 1.0 / 3.0) + pow(1.0 - a, 1.0 / 3.0));
  double z2 = sqrt((3.0 * a * a) + (z1 * z1));
  return 3.0 + z2 - sgn * sqrt((3.0 - z1) * (3.0 + z1 + (2 * z2)));
}

/* event horizon; src/Relphysics.cpp:153-155 */
double orc_kerr_rplus(double a) { return 1 + sqrt(1 - a * a); }

/* u^t of a thin Keplerian disk; src/Relphysics.cpp:28-31 */
double orc_ut_disk(double r, double a) {
  return ((r * sqrt(r) + a) / (sqrt(r) * sqrt(r * r - 3 * r + 2 * a * sqrt(r))));
}

/* src/Relphysics.cpp:127-129 */
static double relat_abberation(double del, double beta) {
  return acos((cos(del) - beta) / (1 - beta * cos(del)));
}

/* src/Relphysics.cpp:158-160 */
static double doppler_factor(double del, double bet) {
  return sqrt(1.0 - bet * bet) / (1.0 + bet * cos(del));
}

/* energy shift lamp post -> disk; src/Relphysics.cpp:163-207 */
double orc_gi_potential_lp(double r, double a, double h, double bet, double del) {
  double ut_d = ((r * sqrt(r) + a) / (sqrt(r) * sqrt(r * r - 3 * r + 2 * a * sqrt(r))));
  double ut_h = sqrt((h * h + a * a) / (h * h - 2 * h + a * a));
  double gi = ut_d / ut_h;
  if (fabs(bet) < 1e-6) {
    return gi;
  }
  double gam = 1.0 / sqrt(1.0 - bet * bet);
  double sgn = (del > M_PI / 2) ? -1.0 : 1.0;
  double delta_eq = h * h - 2 * h + a * a;
  double q2 = (pow(sin(del), 2)) * (pow((h * h + a * a), 2) / delta_eq) - a * a;
  double beta_fac = sqrt(pow((h * h + a * a), 2) - delta_eq * (q2 + a * a));
  beta_fac = gam * (1.0 + sgn * beta_fac / (h * h + a * a) * bet);
  return gi / beta_fac;
}

/* Newtonian lamp-post emissivity; src/Relphysics.cpp:282-288 */
double orc_lp_emissivity_newton(double h, double r) {
  double emis = pow(1. / (pow((r / h), 2) + 1), (3. / 2));
  emis /= (2 * M_PI * h * h);
  return emis;
}

/* g^Gamma * D^2 flux boost; src/Relphysics.cpp:301-311 */
static double fluxboost_source_disk(double rad, double del_emit, double a, double height,
                                    double gamma, double beta) {
  double boost = pow(orc_gi_potential_lp(rad, a, height, beta, del_emit), gamma);
  if (beta > 1e-6) {
    boost *= pow(doppler_factor(del_emit, beta), 2);
  }
  return boost;
}

/* ------------------------------------------------------------------------ */
/* utilities (src/relutility.c)                                              */
/* ------------------------------------------------------------------------ */

double orc_interp_lin_1d(double ifac, double rlo, double rhi) {  /* :24-26 */
  return ifac * rhi + (1.0 - ifac) * rlo;
}

static double interp_log_1d(double ifac, double rlo, double rhi) {  /* :41-43 */
  return exp(ifac * log(rhi) + (1.0 - ifac) * log(rlo));
}

double orc_interp_lin_2d(double i1, double i2, double r11, double r12, double r21,
                         double r22) {  /* :46-51 */
  return (1.0 - i1) * (1.0 - i2) * r11 + (i1) * (1.0 - i2) * r12 + (1.0 - i1) * (i2) * r21 +
         (i1) * (i2) * r22;
}

double orc_interp_lin_2d_float(double i1, double i2, float r11, float r12, float r21,
                               float r22) {  /* :53-58 */
  return (1.0 - i1) * (1.0 - i2) * r11 + (i1) * (1.0 - i2) * r12 + (1.0 - i1) * (i2) * r21 +
         (i1) * (i2) * r22;
}

/* arr ascending: k with arr[k] <= val < arr[k+1]; :155-171 */
int orc_binary_search(const double *arr, int n, double val) {
  if (n <= 1) return -1;
  int klo = 0, khi = n - 1;
  while ((khi - klo) > 1) {
    int k = (khi + klo) / 2;
    if (arr[k] > val) khi = k; else klo = k;
  }
  return klo;
}

int orc_binary_search_float(const float *arr, int n, float val) {  /* :135-151 */
  if (n <= 1) return -1;
  int klo = 0, khi = n - 1;
  while ((khi - klo) > 1) {
    int k = (khi + klo) / 2;
    if (arr[k] > val) khi = k; else klo = k;
  }
  return klo;
}

/* arr descending; :195-211 */
int orc_inv_binary_search(const double *arr, int n, double val) {
  if (n <= 1) return -1;
  int klo = 0, khi = n - 1;
  while ((khi - klo) > 1) {
    int k = (khi + klo) / 2;
    if (arr[k] < val) khi = k; else klo = k;
  }
  return klo;
}

/* half of 2 pi r dr around a descending grid point; :233-244 */
static double trapez_integ_single(const double *re, int ii, int nr) {
  double dr;
  if (ii == 0) {
    dr = 0.5 * (re[ii] - re[ii + 1]);
  } else if (ii == nr - 1) {
    dr = 0.5 * (re[ii - 1] - re[ii]);
  } else {
    dr = 0.5 * (re[ii - 1] - re[ii + 1]);
  }
  return re[ii] * dr * M_PI;
}

static double gstar2ener(double g, double gmin, double gmax, double ener) {  /* :260-262 */
  return (g * (gmax - gmin) + gmin) * ener;
}

/* zone edges; :265-308 */
int orc_get_rzone_grid(double rmin, double rmax, int nzones, double h, double *rgrid) {
  if (nzones == 1) {
    rgrid[0] = rmin;
    rgrid[1] = rmax;
    return 0;
  }
  double r_transition = rmin;
  int indr = 0;
  if (h > rmin) {
    r_transition = h;
    orc_get_log_grid(rgrid, nzones + 1, rmin, rmax);
    indr = orc_binary_search(rgrid, nzones + 1, r_transition);
    r_transition = rgrid[indr];
  }
  if (indr < nzones) {
    double rlo = r_transition, rhi = rmax;
    for (int ii = indr; ii < nzones + 1; ii++) {
      rgrid[ii] = 1.0 * (ii - indr) / (nzones - indr) * (1.0 / rhi - 1.0 / rlo) + 1.0 / rlo;
      rgrid[ii] = fabs(1.0 / rgrid[ii]);
    }
  }
  return 0;
}

/* log grid with n_edges points; :399-405 */
void orc_get_log_grid(double *ener, int n_edges, double emin, double emax) {
  for (int ii = 0; ii < n_edges; ii++) {
    ener[ii] = 1.0 * ii / (n_edges - 1) * (log(emax) - log(emin)) + log(emin);
    ener[ii] = exp(ener[ii]);
  }
}

/* r^{-1/2}-uniform descending fine grid; :666-677 */
void orc_get_fine_radial_grid(double rin, double rout, double *re, int nr) {
  double r1 = 1.0 / sqrt(rout);
  double r2 = 1.0 / sqrt(rin);
  for (int ii = 0; ii < nr; ii++) {
    re[ii] = ((double)(ii)) * (r2 - r1) / (nr - 1) + r1;
    re[ii] = pow(1.0 / (re[ii]), 2);
  }
}

/* radius interpolation factor, log above 75 deg; :407-417 */
static double get_ipol_factor_radius(double rlo, double rhi, double del_inci, double radius) {
  if (del_inci / M_PI * 180.0 <= 75.0) {
    return (radius - rlo) / (rhi - rlo);
  }
  return (log(radius) - log(rlo)) / (log(rhi) - log(rlo));
}

/* float search + float quotient stored in double; :419-423 */
static void get_ipol_factor(const float value, const float *arr, const int n_arr, int *ind,
                            double *ifac) {
  *ind = orc_binary_search_float(arr, n_arr, (float)value);
  *ifac = (value - arr[*ind]) / (arr[*ind + 1] - arr[*ind]);
}

/* bin-overlap rebin; :549-601.  The walk reads ener0[nbins0+1] when an output
 * edge equals ener0[nbins0]; that case has a zero-width overlap and is defined
 * here as 0 (the only in-bounds reading of the reference). */
void orc_rebin_spectrum(const double *ener, double *flu, int nbins, const double *ener0,
                        const double *flu0, int nbins0) {
  int imin = 0, imax = 0;
  for (int ii = 0; ii < nbins; ii++) {
    flu[ii] = 0.0;
    if ((ener0[0] <= ener[ii + 1]) && (ener0[nbins0] >= ener[ii])) {
      while (imin <= nbins0 && ener0[imin] <= ener[ii]) {
        imin++;
      }
      if (imin > 0) imin--;
      while ((ener0[imax] <= ener[ii + 1] && imax < nbins0)) {
        imax++;
      }
      if (imax > 0) imax--;
      if (imin >= nbins0) {  /* zero-width overlap at the top edge */
        continue;
      }
      double elo = ener[ii];
      double ehi = ener[ii + 1];
      if (elo < ener0[imin]) elo = ener0[imin];
      if (ehi > ener0[imax + 1]) ehi = ener0[imax + 1];
      if (imax == imin) {
        flu[ii] = (ehi - elo) / (ener0[imin + 1] - ener0[imin]) * flu0[imin];
      } else {
        double dmin = (ener0[imin + 1] - elo) / (ener0[imin + 1] - ener0[imin]);
        double dmax = (ehi - ener0[imax]) / (ener0[imax + 1] - ener0[imax]);
        flu[ii] += flu0[imin] * dmin + flu0[imax] * dmax;
        for (int jj = imin + 1; jj <= imax - 1; jj++) {
          flu[ii] += flu0[jj];
        }
      }
    }
  }
}

/* mean-flux rebin; :687-725 (table-free known-answer test of the reference) */
void orc_rebin_mean_flux(const double *xn, double *yn, int nn, const double *x0,
                         const double *y0, int n0) {
  int ii = 1;
  for (int in = 0; in < nn; in++) {
    yn[in] = 0.0;
    double xn_m = 0.5 * (xn[in] + xn[in + 1]);
    while (0.5 * (x0[ii - 1] + x0[ii]) < xn_m) {
      ii++;
      if (ii >= n0) break;
    }
    ii--;
    if (ii == 0) ii = 1;
    double x0_m_lo = 0.5 * (x0[ii - 1] + x0[ii]);
    double x0_m_hi = 0.5 * (x0[ii] + x0[ii + 1]);
    if (xn_m > x0_m_lo && xn_m <= x0_m_hi) {
      double ifac_r = (xn_m - x0_m_lo) / (x0_m_hi - x0_m_lo);
      yn[in] = orc_interp_lin_1d(ifac_r, y0[ii - 1], y0[ii]);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* parameters: src/ModelDefinition.cpp:153-334, lmodel order from           */
/* src/modelfiles/lmodel_relxill_public.dat:1-44                             */
/* ------------------------------------------------------------------------ */

int orc_n_params(int model_type) {
  switch (model_type) {
    case ORC_MOD_RELLINE: return 10;
    case ORC_MOD_RELCONV: return 8;
    case ORC_MOD_RELLINELP: return 10;
    case ORC_MOD_RELCONVLP: return 9;
    case ORC_MOD_RELXILL: return 13;
    case ORC_MOD_RELXILLLP: return 14;
    case ORC_MOD_RELXILL_CONV: return 8;  /* relxill-style conv: relconv's parameters */
    case ORC_MOD_RELXILLCP: return 14;
    case ORC_MOD_RELXILLLPCP: return 17;
    case ORC_MOD_RELXILLNS: return 13;
    case ORC_MOD_XILLVER: return 7;
    case ORC_MOD_XILLVERCP: return 8;
    case ORC_MOD_XILLVERNS: return 7;
    default: return -1;
  }
}

/* check_parameter_bounds; src/ModelDefinition.cpp:168-273 */
static int check_parameter_bounds(orc_relparam *p) {
  int status = 0;
  if (p->rin < 0) p->rin = -1.0 * p->rin * orc_kerr_rms(p->a);
  if (p->rout < 0) p->rout = -1.0 * p->rout * orc_kerr_rms(p->a);
  if (p->rbr < 0) p->rbr = -1.0 * p->rbr * orc_kerr_rms(p->a);
  if (p->rout <= p->rin) status = 1;
  double rms = orc_kerr_rms(p->a);
  if (p->rin < rms) p->rin = rms;
  if (p->a > 0.9982) return 1;
  if (p->a < -1) return 1;
  if (p->incl < 3 * M_PI / 180 || p->incl > 87 * M_PI / 180) return 1;
  if (p->rout <= p->rin) return 1;
  if (p->rout > 1000.0) p->rout = 1000.0;
  if (p->emis_type == ORC_EMIS_BKN) {
    if (p->rbr < p->rin) p->rbr = p->rin;
    if (p->rbr > p->rout) p->rbr = p->rout;
  }
  if (p->emis_type == ORC_EMIS_LP) {
    if (p->beta < 0) p->beta = 0.0;
    if (p->beta > 0.99) p->beta = 0.99;
    if (p->height < 0) p->height = -1.0 * p->height * orc_kerr_rplus(p->a);
    if (p->htop < 0) p->htop = -1.0 * p->htop * orc_kerr_rplus(p->a);
    double r_event = orc_kerr_rplus(p->a);
    if ((1.1 * r_event - p->height) > 1e-4) p->height = r_event * 1.1;
  }
  return status;
}

int orc_get_rel_params(int model_type, const double *par, int phys_norm_env, orc_relparam *p) {
  memset(p, 0, sizeof(*p));
  p->model_type = model_type;
  p->num_zones = 1;
  double incl_deg = 0;
  switch (model_type) {
    case ORC_MOD_RELLINE:  /* lineE Index1 Index2 Rbr a Incl Rin Rout z limb */
      p->emis_type = ORC_EMIS_BKN;
      p->lineE = par[0]; p->emis1 = par[1]; p->emis2 = par[2]; p->rbr = par[3];
      p->a = par[4]; incl_deg = par[5]; p->rin = par[6]; p->rout = par[7]; p->z = par[8];
      p->limb = (int)lround(par[9]);
      break;
    case ORC_MOD_RELXILL:  /* Index1 Index2 Rbr a Incl Rin Rout z gamma logxi Afe Ecut refl_frac */
      p->emis_type = ORC_EMIS_BKN;
      p->emis1 = par[0]; p->emis2 = par[1]; p->rbr = par[2]; p->a = par[3]; incl_deg = par[4];
      p->rin = par[5]; p->rout = par[6]; p->z = par[7]; p->gamma = par[8];
      break;
    case ORC_MOD_RELXILLLP:  /* h beta a Incl Rin Rout z gamma logxi Afe Ecut refl_frac
                                switch_returnrad switch_reflfrac_boost */
      p->emis_type = ORC_EMIS_LP;
      p->height = par[0]; p->beta = par[1]; p->a = par[2]; incl_deg = par[3]; p->rin = par[4];
      p->rout = par[5]; p->z = par[6]; p->gamma = par[7];
      p->return_rad = (int)lround(par[12]);
      p->num_zones = 10;  /* N_ZONES, src/common.h:134 (get_num_zones, src/relutility.c:506-544) */
      break;
    case ORC_MOD_RELCONV:  /* Index1 Index2 Rbr a Incl Rin Rout limb */
    case ORC_MOD_RELXILL_CONV:
      p->emis_type = ORC_EMIS_BKN;
      p->emis1 = par[0]; p->emis2 = par[1]; p->rbr = par[2]; p->a = par[3]; incl_deg = par[4];
      p->rin = par[5]; p->rout = par[6]; p->limb = (int)lround(par[7]);
      break;
    case ORC_MOD_RELLINELP:  /* lineE h a Incl Rin Rout z limb gamma switch_returnrad */
      p->emis_type = ORC_EMIS_LP;
      p->lineE = par[0]; p->height = par[1]; p->a = par[2]; incl_deg = par[3]; p->rin = par[4];
      p->rout = par[5]; p->z = par[6]; p->limb = (int)lround(par[7]); p->gamma = par[8];
      p->return_rad = (int)lround(par[9]);
      break;
    case ORC_MOD_RELCONVLP:  /* h beta a Incl Rin Rout limb gamma switch_returnrad */
      p->emis_type = ORC_EMIS_LP;
      p->height = par[0]; p->beta = par[1]; p->a = par[2]; incl_deg = par[3]; p->rin = par[4];
      p->rout = par[5]; p->limb = (int)lround(par[6]); p->gamma = par[7];
      p->return_rad = (int)lround(par[8]);
      break;
    default:
      return 1;
  }
  p->incl = incl_deg * M_PI / 180;  /* src/ModelDefinition.cpp:293 */
  /* do_renorm_model: src/relutility.c:603-623 */
  if (model_type < 0) {
    p->do_renorm = (p->emis_type == ORC_EMIS_LP || phys_norm_env) ? 0 : 1;
  } else {
    p->do_renorm = phys_norm_env ? 0 : 1;
  }
  if (check_parameter_bounds(p)) return 1;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* system parameters                                                         */
/* ------------------------------------------------------------------------ */

orc_syspar *orc_syspar_new(int nr, int ng) {
  orc_syspar *s = (orc_syspar *)calloc(1, sizeof(orc_syspar));
  s->nr = nr;
  s->ng = ng;
  s->re = (double *)calloc(nr, sizeof(double));
  s->gmin = (double *)calloc(nr, sizeof(double));
  s->gmax = (double *)calloc(nr, sizeof(double));
  s->emis = (double *)calloc(nr, sizeof(double));
  s->del_emit = (double *)calloc(nr, sizeof(double));
  s->del_inc = (double *)calloc(nr, sizeof(double));
  s->trff = (double *)calloc((size_t)nr * ng * 2, sizeof(double));
  s->cosne = (double *)calloc((size_t)nr * ng * 2, sizeof(double));
  s->gstar = (double *)calloc(ng, sizeof(double));
  s->d_gstar = (double *)calloc(ng, sizeof(double));
  /* g* grid and bin widths; src/Relprofile.cpp:104-116 */
  for (int ii = 0; ii < ng; ii++) {
    s->gstar[ii] = GFAC_H + (1.0 - 2 * GFAC_H) / (ng - 1) * ((float)(ii));
  }
  for (int ii = 0; ii < ng; ii++) {
    if ((ii == 0) || (ii == (ng - 1))) {
      s->d_gstar[ii] = 0.5 * (s->gstar[1] - s->gstar[0]) + GFAC_H;
    } else {
      s->d_gstar[ii] = s->gstar[1] - s->gstar[0];
    }
  }
  return s;
}

void orc_syspar_free(orc_syspar *s) {
  if (!s) return;
  free(s->re); free(s->gmin); free(s->gmax); free(s->emis); free(s->del_emit);
  free(s->del_inc); free(s->trff); free(s->cosne); free(s->gstar); free(s->d_gstar);
  free(s);
}

#define NODE(tab, ia, im) ((size_t)(ia) * (tab)->n_mu0 + (im))

/* interpol_relTable; src/Relprofile.cpp:141-303 (with interpol_a_mu0 :39-80)
 * Pinned by closed forms (tests/test_oracle_kats.py): the (a, mu0) bilinear
 * node on an affine field, the spin-lerped node radii and the node -> fine
 * mapping on a radius-affine field; the ISCO and 1000 rg snaps (:195-200,
 * :235-238) by the ramps they put into the innermost and outermost node
 * intervals of a radius-valued table. */
int orc_interpol_reltable(const orc_reltable *tab, double a, double incl, double rin,
                          double rout, orc_syspar *s) {
  const int nr = tab->n_r, ng = tab->n_g;
  double rms = orc_kerr_rms(a);
  if (!(rout > rin) || !(rin >= rms)) return 10;
  double mu0 = cos(incl);

  int ind_a = orc_binary_search_float(tab->a, tab->n_a, (float)a);
  int ind_mu0 = orc_binary_search_float(tab->mu0, tab->n_mu0, (float)mu0);
  float ifac_a = ((float)a - tab->a[ind_a]) / (tab->a[ind_a + 1] - tab->a[ind_a]);
  float ifac_mu0 = ((float)mu0 - tab->mu0[ind_mu0]) / (tab->mu0[ind_mu0 + 1] - tab->mu0[ind_mu0]);
  if (!(ifac_mu0 >= 0 && ifac_mu0 <= 1 && ifac_a >= 0 && ifac_a <= 1)) return 11;

  const size_t n00 = NODE(tab, ind_a, ind_mu0), n10 = NODE(tab, ind_a + 1, ind_mu0);
  const size_t n01 = NODE(tab, ind_a, ind_mu0 + 1), n11 = NODE(tab, ind_a + 1, ind_mu0 + 1);

  double *tre = (double *)malloc(sizeof(double) * nr);
  double *tgmin = (double *)malloc(sizeof(double) * nr);
  double *tgmax = (double *)malloc(sizeof(double) * nr);
  double *ttrff = (double *)malloc(sizeof(double) * nr * ng * 2);
  double *tcosne = (double *)malloc(sizeof(double) * nr * ng * 2);
  int status = 0;

  /* radius of the (a, mu0) node: lerp in a only (:190-193) */
  for (int ii = 0; ii < nr; ii++) {
    tre[ii] = orc_interp_lin_1d(ifac_a, tab->r[n00 * nr + ii], tab->r[n10 * nr + ii]);
  }
  /* innermost radius snapped to the ISCO (:195-200) */
  if ((tre[nr - 1] > orc_kerr_rms(a)) && ((tre[nr - 1] - orc_kerr_rms(a)) / tre[nr - 1] < 1e-3)) {
    tre[nr - 1] = orc_kerr_rms(a);
  }
  int ind_rmin = orc_inv_binary_search(tre, nr, rin);
  int ind_rmax = orc_inv_binary_search(tre, nr, rout);

  for (int ii = 0; ii < nr; ii++) {
    if (ii <= ind_rmin || ii >= ind_rmax + 1) {  /* always true (:211) */
      const size_t o = (size_t)ii;
      tgmin[ii] = orc_interp_lin_2d_float(ifac_a, ifac_mu0, tab->gmin[n00 * nr + o],
                                          tab->gmin[n10 * nr + o], tab->gmin[n01 * nr + o],
                                          tab->gmin[n11 * nr + o]);
      tgmax[ii] = orc_interp_lin_2d_float(ifac_a, ifac_mu0, tab->gmax[n00 * nr + o],
                                          tab->gmax[n10 * nr + o], tab->gmax[n01 * nr + o],
                                          tab->gmax[n11 * nr + o]);
      for (int jj = 0; jj < ng; jj++) {
        const size_t q = (size_t)ii * ng + jj;
        const size_t s00 = n00 * nr * ng + q, s10 = n10 * nr * ng + q;
        const size_t s01 = n01 * nr * ng + q, s11 = n11 * nr * ng + q;
        ttrff[q * 2 + 0] = orc_interp_lin_2d_float(ifac_a, ifac_mu0, tab->trff1[s00],
                                                   tab->trff1[s10], tab->trff1[s01], tab->trff1[s11]);
        ttrff[q * 2 + 1] = orc_interp_lin_2d_float(ifac_a, ifac_mu0, tab->trff2[s00],
                                                   tab->trff2[s10], tab->trff2[s01], tab->trff2[s11]);
        tcosne[q * 2 + 0] = orc_interp_lin_2d_float(ifac_a, ifac_mu0, tab->cosne1[s00],
                                                    tab->cosne1[s10], tab->cosne1[s01], tab->cosne1[s11]);
        tcosne[q * 2 + 1] = orc_interp_lin_2d_float(ifac_a, ifac_mu0, tab->cosne2[s00],
                                                    tab->cosne2[s10], tab->cosne2[s01], tab->cosne2[s11]);
      }
    } else {
      tgmin[ii] = 0.0;
      tgmax[ii] = 0.0;
      for (int jj = 0; jj < ng * 2; jj++) {
        ttrff[(size_t)ii * ng * 2 + jj] = 0.0;
        tcosne[(size_t)ii * ng * 2 + jj] = 0.0;
      }
    }
  }

  orc_get_fine_radial_grid(rin, rout, s->re, s->nr);

  /* outermost node radius snapped to 1000 rg (:235-238) */
  const double rmax = 1000.0;
  if (tre[ind_rmax] < rmax && tre[ind_rmax] * 1.01 > rmax) {
    tre[ind_rmax] = rmax;
  }
  if (!(ind_rmin > 0) || !(tre[ind_rmin + 1] <= rin) || !(tre[ind_rmin] >= rin) ||
      !(tre[ind_rmax + 1] <= rout) || !(tre[ind_rmax] >= rout) || !(ind_rmax <= ind_rmin) ||
      !(rout <= 1000.0)) {
    status = 12;
    goto done;
  }

  /* linear map of the node onto the fine grid (:249-300) */
  {
    int ind_tabr = ind_rmin;
    for (int ii = s->nr - 1; ii >= 0; ii--) {
      while ((s->re[ii] >= tre[ind_tabr])) {
        ind_tabr--;
        if (ind_tabr < 0) {
          if (s->re[ii] - RELTABLE_MAX_R <= 1e-6) {
            ind_tabr = 0;
            break;
          } else {
            status = 13;
            goto done;
          }
        }
      }
      double ifac_r = (s->re[ii] - tre[ind_tabr + 1]) / (tre[ind_tabr] - tre[ind_tabr + 1]);
      if (ifac_r > 1.0 && ind_tabr > 0) {
        status = 14;
        goto done;
      }
      for (int jj = 0; jj < ng; jj++) {
        for (int kk = 0; kk < 2; kk++) {
          const size_t lo = ((size_t)(ind_tabr + 1) * ng + jj) * 2 + kk;
          const size_t hi = ((size_t)(ind_tabr)*ng + jj) * 2 + kk;
          s->trff[((size_t)ii * ng + jj) * 2 + kk] = orc_interp_lin_1d(ifac_r, ttrff[lo], ttrff[hi]);
          s->cosne[((size_t)ii * ng + jj) * 2 + kk] = orc_interp_lin_1d(ifac_r, tcosne[lo], tcosne[hi]);
        }
      }
      s->gmin[ii] = orc_interp_lin_1d(ifac_r, tgmin[ind_tabr + 1], tgmin[ind_tabr]);
      s->gmax[ii] = orc_interp_lin_1d(ifac_r, tgmax[ind_tabr + 1], tgmax[ind_tabr]);
    }
  }

done:
  free(tre); free(tgmin); free(tgmax); free(ttrff); free(tcosne);
  return status;
}

/* ------------------------------------------------------------------------ */
/* emissivity (src/Rellp.cpp)                                                */
/* ------------------------------------------------------------------------ */

/* norm_emis_profile; src/Rellp.cpp:91-109 (descending grid) */
static void norm_emis_profile(const double *re, int nr, double *emis) {
  double integ_area = 0.0;
  for (int ii = 0; ii < nr; ii++) {
    double delta_area = trapez_integ_single(re, ii, nr) * 2;
    integ_area += emis[ii] * delta_area;
  }
  for (int ii = 0; ii < nr; ii++) {
    emis[ii] /= integ_area;
  }
}

/* get_emis_bkn; src/Rellp.cpp:421-437 */
static void get_emis_bkn(double *emis, const double *re, int nr, double index1, double index2,
                         double rbr) {
  for (int ii = 0; ii < nr; ii++) {
    double alpha = index1;
    if (re[ii] > rbr) alpha = index2;
    emis[ii] = pow(re[ii] / rbr, -alpha);
  }
  norm_emis_profile(re, nr, emis);
}

/* calc_refl_frac; src/Rellp.cpp:36-89 */
static int calc_refl_frac(const orc_syspar *s, double rin, double rout, double del_emit_ad_max,
                          double beta, orc_reflfrac *str) {
  double del_bh = s->del_emit[orc_inv_binary_search(s->re, s->nr, rin)];
  double del_ad = s->del_emit[orc_inv_binary_search(s->re, s->nr, rout)];
  if (del_emit_ad_max < M_PI / 2.0) del_emit_ad_max = M_PI / 2.0;
  if (beta > 1e-6) {
    del_bh = relat_abberation(del_bh, -1. * beta);
    del_ad = relat_abberation(del_ad, -1. * beta);
  }
  str->f_bh = 0.5 * (1.0 - cos(del_bh));
  str->f_ad = 0.5 * (cos(del_bh) - cos(del_ad));
  str->f_inf_rest = 0.5 * (1.0 + cos(del_emit_ad_max));
  if (beta > 1e-6) {
    str->f_inf = 0.5 * (1.0 + cos(relat_abberation(del_emit_ad_max, -1. * beta)));
  } else {
    str->f_inf = str->f_inf_rest;
  }
  if (!(str->f_inf + str->f_ad < 1.0)) return 20;
  str->refl_frac = str->f_ad / str->f_inf;
  return 0;
}

/* lamp-post point source: interpol_lptable (:219-247),
 * interpol_emisprofile_spin_height (:190-210), rebin_emisprofile_on_radial_grid
 * (:120-175), calc_refl_frac (:36-89), apply_emis_fluxboost_source_disk
 * (:249-253), all from calc_emis_jet_point_source (:266-285).
 * Pinned by closed forms: the 75 deg lin/log switch of the radius factor on a
 * table whose fields are exact under one factor on each side of the switch;
 * the boost g^gamma by the u^t ratio. */
static int emis_lamp_post(const orc_lptable *tab, const orc_relparam *p, orc_syspar *s) {
  const int nh = tab->n_h, ntr = tab->n_r;
  int ind_a;
  double ifac_a;
  get_ipol_factor((float)p->a, tab->a, tab->n_a, &ind_a, &ifac_a);
  const int ia[2] = {ind_a, ind_a + 1};

  double *jet_rad = (double *)malloc(sizeof(double) * ntr);
  double *t_emis = (double *)malloc(sizeof(double) * ntr);
  double *t_del = (double *)malloc(sizeof(double) * ntr);
  double *t_dinc = (double *)malloc(sizeof(double) * ntr);
  int status = 0;
  for (int ii = 0; ii < ntr; ii++) {
    jet_rad[ii] = orc_interp_lin_1d(ifac_a, tab->rad[(size_t)ia[0] * ntr + ii],
                                    tab->rad[(size_t)ia[1] * ntr + ii]);
  }
  int ind_h[2];
  double ifac_h[2];
  for (int k = 0; k < 2; k++) {
    get_ipol_factor((float)p->height, tab->h + (size_t)ia[k] * nh, nh, &ind_h[k], &ifac_h[k]);
  }
  for (int ii = 0; ii < ntr; ii++) {
    const size_t l0 = ((size_t)ia[0] * nh + ind_h[0]) * ntr + ii;
    const size_t h0 = ((size_t)ia[0] * nh + ind_h[0] + 1) * ntr + ii;
    const size_t l1 = ((size_t)ia[1] * nh + ind_h[1]) * ntr + ii;
    const size_t h1 = ((size_t)ia[1] * nh + ind_h[1] + 1) * ntr + ii;
    t_emis[ii] = (1.0 - ifac_a) * orc_interp_lin_1d(ifac_h[0], tab->intens[l0], tab->intens[h0]) +
                 (ifac_a)*orc_interp_lin_1d(ifac_h[1], tab->intens[l1], tab->intens[h1]);
    t_del[ii] = (1.0 - ifac_a) * orc_interp_lin_1d(ifac_h[0], tab->del[l0], tab->del[h0]) +
                (ifac_a)*orc_interp_lin_1d(ifac_h[1], tab->del[l1], tab->del[h1]);
    t_dinc[ii] = (1.0 - ifac_a) * orc_interp_lin_1d(ifac_h[0], tab->del_inc[l0], tab->del_inc[h0]) +
                 (ifac_a)*orc_interp_lin_1d(ifac_h[1], tab->del_inc[l1], tab->del_inc[h1]);
  }

  /* table (ascending radii) -> fine grid (descending) */
  {
    const double *re = s->re;
    const int nr = s->nr;
    int kk = orc_binary_search(jet_rad, ntr, re[nr - 1]);
    if (kk < 0 || kk >= ntr - 1) {
      status = 21;
      goto done;
    }
    for (int ii = nr - 1; ii >= 0; ii--) {
      while ((re[ii] >= jet_rad[kk + 1])) {
        kk++;
        if (kk >= ntr - 1) {
          if (re[ii] - RELTABLE_MAX_R <= 1e-6) {
            kk = ntr - 2;
            break;
          } else {
            status = 22;
            goto done;
          }
        }
      }
      const double inter_r = get_ipol_factor_radius(jet_rad[kk], jet_rad[kk + 1], t_del[kk], re[ii]);
      s->emis[ii] = interp_log_1d(inter_r, t_emis[kk], t_emis[kk + 1]);
      s->del_emit[ii] = orc_interp_lin_1d(inter_r, t_del[kk], t_del[kk + 1]);
      s->del_inc[ii] = orc_interp_lin_1d(inter_r, t_dinc[kk], t_dinc[kk + 1]);
    }
  }

  status = calc_refl_frac(s, p->rin, p->rout, t_del[ntr - 1], p->beta, &s->refl);
  if (status) goto done;

  for (int ii = 0; ii < s->nr; ii++) {
    s->emis[ii] *= fluxboost_source_disk(s->re[ii], s->del_emit[ii], p->a, p->height, p->gamma, p->beta);
  }

done:
  free(jet_rad); free(t_emis); free(t_del); free(t_dinc);
  return status;
}

/* ---- returning radiation (src/Relreturn_Corona.cpp, src/Relreturn_Table.cpp) ---- */

/* inv_rebin_mean (src/relutility.c:636-663): y0 on the descending x0,
 * interpolated at the ascending xn */
static void inv_rebin_mean(const double *x0, const double *y0, int n0, const double *xn, double *yn, int nn) {
  int in = nn - 1;
  for (int ii = 0; ii < n0 - 1; ii++) {
    if (x0[ii] > xn[in] && x0[ii + 1] <= xn[in]) {
      const double ifac_r = (xn[in] - x0[ii + 1]) / (x0[ii] - x0[ii + 1]);
      yn[in] = orc_interp_lin_1d(ifac_r, y0[ii + 1], y0[ii]);
      in--;
      if (in < 0) break;
    }
  }
}



int orc_rrad_emis(const orc_rradtable *rt, const orc_relparam *p, const double *re, const double *emis0, int nr,
                  double *emis_ret) {
  const double a = p->a;
  /* determine_rlo_rhi (:201-213): the fine grid's ends */
  const double rin = re[nr - 1], rout = re[0];
  /* select_spinIndexForTable (src/Relreturn_Table.cpp:343-357): the first
   * tabulated spin at or above a */
  int k = orc_binary_search(rt->spin, rt->n_spin, a);
  if (k < 0) k = 0;
  if (rt->spin[k] < a) k++;
  if (k >= rt->n_spin) return 35;
  const int nt = rt->n_rad, ng = rt->n_g;
  const double *rlo_t = rt->rlo + (size_t)k * nt, *rhi_t = rt->rhi + (size_t)k * nt;
  /* allocate_radial_grid (:359-404): the table zones covering [rin, rout],
   * the outer ends replaced by rin and rout */
  int klo = orc_binary_search(rlo_t, nt, rin);
  int khi = orc_binary_search(rhi_t, nt, rout);
  if (fabs(rout - rhi_t[nt - 1]) < 1e-6) khi = nt - 1; else khi++;
  if (fabs(rin - rlo_t[0]) < 1e-6) klo = 0;
  const int n = (khi + 1) - klo;
  if (n < 2 || n > nt || klo < 0) return 36;
  if (!(rin >= rlo_t[klo] - 1e-4)) return 37;
  double *rlo = (double *)malloc(sizeof(double) * n), *rhi = (double *)malloc(sizeof(double) * n);
  double *rad = (double *)malloc(sizeof(double) * n), *emis_in = (double *)malloc(sizeof(double) * n);
  double *ret = (double *)malloc(sizeof(double) * n), *tfr = (double *)malloc(sizeof(double) * n * n);
  double *term = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) {
    rlo[i] = rlo_t[klo + i];
    rhi[i] = rhi_t[klo + i];
  }
  rlo[0] = rin;
  rhi[n - 1] = rout;
  for (int i = 0; i < n; i++) rad[i] = 0.5 * (rlo[i] + rhi[i]);
  /* get_interpolated_tfr (:437-480) with the ring-area corrections of the end
   * zones (get_ring_area_correction_factor, :415-435) */
  const double *tf = rt->tf_r + (size_t)k * nt * nt;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) tfr[i * n + j] = tf[(size_t)(klo + i) * nt + (klo + j)];
  double corr[2];
  for (int e = 0; e < 2; e++) {
    const int i = e == 0 ? 0 : n - 1;
    double rl = rlo_t[klo + i];
    if (klo + i == 0 && rl > orc_kerr_rms(a)) rl = orc_kerr_rms(a);
    const double rh = rhi_t[klo + i];
    const double area_table = 0.5 * (rl + rh) * (rh - rl);
    const double area_model = 0.5 * (rlo[i] + rhi[i]) * (rhi[i] - rlo[i]);
    corr[e] = area_model / area_table;
  }
  for (int i = 0; i < n - 1; i++) tfr[i * n + 0] *= corr[0];
  for (int i = 0; i < n - 2; i++) tfr[i * n + n - 1] *= corr[1];
  /* the direct emissivity at the zone centres (inv_rebin_mean, src/relutility.c:636-663) */
  inv_rebin_mean(re, emis0, nr, rad, emis_in, n);
  /* relxilllp's correction factors of the zone each table zone's centre
   * falls in (rebin_corrfactors_to_rradtable_grid with rebin_to_grid,
   * src/Relreturn_Corona.cpp:248-289; a centre beyond the last edge takes the
   * last zone, where the reference would read one past its array) */
  double *cf_flux = (double *)malloc(sizeof(double) * n), *cf_g = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) {
    cf_flux[i] = 1.0;
    cf_g[i] = 1.0;
    if (p->rr_corr_n > 0) {
      const int n0 = p->rr_corr_n;
      int ind = orc_binary_search(p->rr_corr_rgrid, n0 + 1, rad[i]);
      if (rad[i] < p->rr_corr_rgrid[0]) ind = 0;
      else if (rad[i] > p->rr_corr_rgrid[n0] || ind > n0 - 1) ind = n0 - 1;
      cf_flux[i] = p->rr_corr_flux[ind];
      cf_g[i] = p->rr_corr_gshift[ind];
    }
  }
  /* calc_rrad_emis_corona (src/Relreturn_Corona.cpp:119-151) with
   * calc_rrad_emis_zone (:87-108): sum over g of frac_g g^(Gamma-1), or with
   * a correction factor of the emitting zone off 1 by more than 1e-3 of
   * frac_g times its corrected boost over g */
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) {
      const size_t q = (size_t)k * nt * nt + (size_t)(klo + i) * nt + (klo + j);
      const double gmin = rt->gmin[q], gmax = rt->gmax[q];
      double ez = 0.0;
      for (int jj = 0; jj < ng; jj++) {
        const double g = (jj + 0.5) / ng * (gmax - gmin) + gmin;  /* get_gfac_grid, src/Relreturn_Datastruct.cpp:51-59 */
        double v = rt->frac_g[q * ng + jj];
        if (fabs(cf_g[j] - 1) > 1e-3) v *= orc_rrad_gshift_boost(cf_g[j], g, p->gamma) / g;
        else if (fabs(g - 1) > 1e-3) v *= pow(g, p->gamma - 1);
        ez += v;
      }
      term[j] = ez * tfr[i * n + j] * emis_in[j];
    }
    double sum = 0.0;
    for (int j = 0; j < n; j++) sum += term[j];
    ret[i] = sum * cf_flux[i];
  }
  free(cf_flux); free(cf_g);
  /* back to the fine grid (rebin_emisprofile_on_radial_grid, src/Rellp.cpp:120-175;
   * the profile's emission angles are 0, so the radius factor is linear and
   * the emissivity log-interpolated) */
  int status = 0;
  int kk = orc_binary_search(rad, n, re[nr - 1]);
  if (kk < 0 || kk >= n - 1) status = 38;
  for (int ii = nr - 1; ii >= 0 && status == 0; ii--) {
    while (re[ii] >= rad[kk + 1]) {
      kk++;
      if (kk >= n - 1) {
        if (re[ii] - RELTABLE_MAX_R <= 1e-6) {
          kk = n - 2;
          break;
        }
        status = 39;
        break;
      }
    }
    const double inter_r = get_ipol_factor_radius(rad[kk], rad[kk + 1], 0.0, re[ii]);
    emis_ret[ii] = interp_log_1d(inter_r, ret[kk], ret[kk + 1]);
  }
  free(rlo); free(rhi); free(rad); free(emis_in); free(ret); free(tfr); free(term);
  return status;
}

/* calc_emis_profile; src/Rellp.cpp:518-564 (BKN and LP point source) */
int orc_calc_emis_profile(const orc_lptable *lp, const orc_relparam *p, orc_syspar *s) {
  if (p->emis_type == ORC_EMIS_BKN) {
    get_emis_bkn(s->emis, s->re, s->nr, p->emis1, p->emis2, p->rbr);
    for (int ii = 0; ii < s->nr; ii++) {
      s->del_emit[ii] = -1.0;
      s->del_inc[ii] = -1.0;
    }
    return 0;
  }
  if (p->emis_type == ORC_EMIS_LP) {
    if (!lp) return 30;
    /* point source unless htop describes an extended jet (:287-294) */
    if (!((fabs(p->htop) <= 1.0) || (p->htop - 1e-3 <= p->height))) return 31;
    int st = emis_lamp_post(lp, p, s);
    if (st == 0 && p->return_rad != 0) {
      /* add_returnrad_emis (src/Rellp.cpp:483-507) */
      if (!lp->rrad) return 32;
      if (!(p->return_rad == 1 || p->return_rad == -1 || p->return_rad == 2)) return 34;
      double *er = (double *)malloc(sizeof(double) * s->nr);
      st = orc_rrad_emis(lp->rrad, p, s->re, s->emis, s->nr, er);
      if (st == 0) {
        for (int ii = 0; ii < s->nr; ii++) s->emis[ii] = (p->return_rad == 1) ? s->emis[ii] + er[ii] : er[ii];
      }
      free(er);
    }
    return st;
  }
  return 33;
}

orc_syspar *orc_get_system_parameters(const orc_reltable *tab, const orc_lptable *lp,
                                      const orc_relparam *p, int nr_fine, int *status) {
  orc_syspar *s = orc_syspar_new(nr_fine, tab->n_g);
  *status = orc_interpol_reltable(tab, p->a, p->incl, p->rin, p->rout, s);
  if (*status == 0) {
    s->limb_law = p->limb;
    *status = orc_calc_emis_profile(lp, p, s);
  }
  if (*status) {
    orc_syspar_free(s);
    return NULL;
  }
  return s;
}

/* ------------------------------------------------------------------------ */
/* relline integration (src/Relprofile.cpp:489-958)                          */
/* ------------------------------------------------------------------------ */

typedef struct {
  double re, gmin, gmax, del_g, emis;
  int limb_law, ng;
  const double *trff, *cosne, *gstar; /* trff/cosne: [ng][2] of this radius */
  orc_stats *st;
} relb_str;

/* relb_func; src/Relprofile.cpp:489-521 */
static double relb_func(double eg, int k, const relb_str *str) {
  if (str->st) str->st->relb_func_calls++;
  double egstar = (eg - str->gmin) * str->del_g;
  int ind = orc_binary_search(str->gstar, str->ng, egstar);
  double inte = (egstar - str->gstar[ind]) / (str->gstar[ind + 1] - str->gstar[ind]);
  double inte1 = 1.0 - inte;
  double ftrf = inte * str->trff[ind * 2 + k] + inte1 * str->trff[(ind + 1) * 2 + k];
  double val = pow(eg, 3) / ((str->gmax - str->gmin) * sqrt(egstar - egstar * egstar)) * ftrf * str->emis;
  if (str->limb_law == 0) return val;
  double fmu0 = inte * str->cosne[ind * 2 + k] + inte1 * str->cosne[(ind + 1) * 2 + k];
  double limb = 1.0;
  if (str->limb_law == 1) {
    limb = (1.0 + 2.06 * fmu0);  /* Laor (1991) */
  } else if (str->limb_law == 2) {
    limb = log(1.0 + 1.0 / fmu0);  /* Haardt (1993) */
  }
  return val * limb;
}

/* romberg_integration; src/Relprofile.cpp:524-579 */
static double romberg_integration(double a, double b, int k, const relb_str *str) {
  const double prec = 0.02;
  double obtprec = 1.0;
  const int itermin = 0;
  int itermax = 5;
  enum { maxiter = 6 };
  double t[maxiter + 1][maxiter + 1];
  int niter = 0;
  if (str->st) str->st->romberg_calls++;
  double r = relb_func(a, k, str);
  double fb = relb_func(b, k, str);
  double ta = (r + fb) / 2.0;
  double pas = b - a;
  t[0][0] = ta * pas;
  while ((niter < itermin) || ((obtprec > prec) && (niter <= itermax))) {
    niter++;
    if (str->st) str->st->romberg_levels++;
    pas = pas / 2.0;
    double s = ta;
    for (int ii = 1; ii <= pow(2, niter) - 1; ii++) {
      s += relb_func(a + pas * ii, k, str);
    }
    t[0][niter] = s * pas;
    r = 1.0;
    for (int ii = 1; ii <= niter; ii++) {
      r *= 4.0;
      int jj = niter - ii;
      t[ii][jj] = (r * t[ii - 1][jj + 1] - t[ii - 1][jj]) / (r - 1.0);
    }
    obtprec = fabs(t[niter][0] - t[niter - 1][0]) / t[niter][0];
  }
  return t[niter][0];
}

/* int_edge: analytic 1/sqrt endpoint integral; src/Relprofile.cpp:585-621 */
static double int_edge(double blo, double bhi, double h, const relb_str *str, double line_energy) {
  double hex, lo, hi;
  if (str->st) str->st->edge_calls++;
  if (blo <= 0.5) {
    hex = h;
    lo = blo;
    hi = bhi;
  } else {
    hex = 1.0 - h;
    lo = 1.0 - bhi;
    hi = 1.0 - blo;
  }
  double norm = 0.0;
  for (int k = 0; k < 2; k++) {
    norm = norm + relb_func(gstar2ener(hex, str->gmin, str->gmax, line_energy), k, str);
  }
  norm = norm * sqrt(h);
  return 2 * norm * (sqrt(hi) - sqrt(lo)) * line_energy * (str->gmax - str->gmin);
}

/* int_romb: Romberg above 0.95 E_line, midpoint below; :628-647 */
static double int_romb(double lo, double hi, const relb_str *str, double line_energy) {
  double flu = 0.0;
  if (lo >= line_energy * 0.95) {
    for (int k = 0; k < 2; k++) flu += romberg_integration(lo, hi, k, str);
  } else {
    if (str->st) str->st->midpoint_calls++;
    for (int k = 0; k < 2; k++) flu += relb_func((hi + lo) / 2.0, k, str) * (hi - lo);
  }
  return flu;
}

/* integ_relline_bin; src/Relprofile.cpp:650-726 */
static double integ_relline_bin(const relb_str *str, double rlo0, double rhi0) {
  const double line_ener = 1.0;
  double flu = 0.0;
  if (str->st) str->st->bin_integrals++;
  double gblo = (rlo0 / line_ener - str->gmin) * str->del_g;
  if (gblo < 0.0) gblo = 0.0; else if (gblo > 1.0) gblo = 1.0;
  double gbhi = (rhi0 / line_ener - str->gmin) * str->del_g;
  if (gbhi < 0.0) gbhi = 0.0; else if (gbhi > 1.0) gbhi = 1.0;
  if (gbhi == 0) return 0.0;
  double rlo = rlo0, rhi = rhi0, hlo, hhi;
  if (gblo <= GFAC_H) {
    hlo = gblo;
    hhi = GFAC_H;
    rlo = gstar2ener(GFAC_H, str->gmin, str->gmax, line_ener);
    if (gbhi <= GFAC_H) {
      hhi = gbhi;
      rlo = -1.0;
    }
    flu = flu + int_edge(hlo, hhi, GFAC_H, str, line_ener);
  }
  if (gbhi >= (1.0 - GFAC_H)) {
    hhi = gbhi;
    hlo = 1.0 - GFAC_H;
    rhi = gstar2ener(1 - GFAC_H, str->gmin, str->gmax, line_ener);
    if (gblo >= (1.0 - GFAC_H)) {
      hlo = gblo;
      rhi = -1.0;
    }
    flu = flu + int_edge(hlo, hhi, GFAC_H, str, line_ener);
  }
  if ((rhi >= 0) && (rlo >= 0)) {
    flu = flu + int_romb(rlo, rhi, str, line_ener);
  }
  return flu;
}

/* calc_relline_profile; src/Relprofile.cpp:835-958 */
int orc_calc_relline_profile(const orc_syspar *s, const double *ener, int n_ener,
                             const double *rgrid, int nzones, double *flux, double *dist,
                             int n_incl, orc_stats *st) {
  const double line_ener = 1.0;
  int status = 0;
  memset(flux, 0, sizeof(double) * (size_t)nzones * n_ener);
  if (dist) memset(dist, 0, sizeof(double) * (size_t)nzones * n_incl);
  relb_str str;
  str.gstar = s->gstar;
  str.ng = s->ng;
  str.st = st;
  for (int ii = 0; ii < s->nr; ii++) {
    double egmin = s->gmin[ii] * line_ener;
    double egmax = s->gmax[ii] * line_ener;
    if ((egmax > ener[0]) && (egmin < ener[n_ener])) {
      if (egmin < ener[0]) egmin = ener[0];
      if (egmax > ener[n_ener]) egmax = ener[n_ener];
      int ielo = orc_binary_search(ener, n_ener + 1, egmin);
      int iehi = orc_binary_search(ener, n_ener + 1, egmax);
      int izone = orc_binary_search(rgrid, nzones + 1, s->re[ii]);
      /* set_str_relbf (:728-746) */
      str.re = s->re[ii];
      str.gmin = s->gmin[ii];
      str.gmax = s->gmax[ii];
      str.del_g = 1. / (s->gmax[ii] - s->gmin[ii]);
      str.emis = s->emis[ii];
      str.trff = s->trff + (size_t)ii * s->ng * 2;
      str.cosne = s->cosne + (size_t)ii * s->ng * 2;
      str.limb_law = s->limb_law;
      double weight = trapez_integ_single(s->re, ii, s->nr) / 2;
      for (int jj = ielo; jj <= iehi; jj++) {
        double tmp_var = integ_relline_bin(&str, ener[jj], ener[jj + 1]);
        flux[(size_t)izone * n_ener + jj] += tmp_var * weight;
      }
      if (dist) {  /* angular distribution (:908-938) */
        for (int jj = 0; jj < s->ng; jj++) {
          double g = s->gstar[jj] * (str.gmax - str.gmin) + str.gmin;
          for (int kk = 0; kk < 2; kk++) {
            int imu = ((int)(n_incl * (1 - str.cosne[jj * 2 + kk]) + 1)) - 1;
            if (imu < 0 || imu >= n_incl) return 40;
            double tmp = str.re * pow(2 * M_PI * g * str.re, 2) /
                         sqrt(s->gstar[jj] - s->gstar[jj] * s->gstar[jj]) * str.trff[jj * 2 + kk] *
                         str.emis * weight * s->d_gstar[jj];
            if (tmp != tmp) status = 41;
            dist[(size_t)izone * n_incl + imu] += tmp;
          }
        }
      }
    }
  }
  return status;
}

/* renorm_relline_profile; src/Relprofile.cpp:749-796 */
void orc_renorm_relline_profile(const double *ener, int n_ener, int nzones, double *flux,
                                double *dist, int n_incl, const orc_relparam *p) {
  double sum = 0.0;
  for (int ii = 0; ii < nzones; ii++) {
    for (int jj = 0; jj < n_ener; jj++) {
      flux[(size_t)ii * n_ener + jj] /= 0.5 * (ener[jj] + ener[jj + 1]);
      sum += flux[(size_t)ii * n_ener + jj];
    }
  }
  if (p->do_renorm) {
    double relline_norm = 1;
    if (p->model_type < 0 && p->emis_type == ORC_EMIS_BKN) {
      /* norm_factor_semi_infinite_slab, src/Xillspec.cpp:528-530 */
      relline_norm = 0.5 * cos((p->incl * 180.0 / M_PI) * M_PI / 180);
    }
    for (int ii = 0; ii < nzones; ii++) {
      for (int jj = 0; jj < n_ener; jj++) {
        flux[(size_t)ii * n_ener + jj] *= relline_norm / sum;
      }
    }
  }
  if (dist) {
    for (int ii = 0; ii < nzones; ii++) {
      double dsum = 0.0;
      for (int jj = 0; jj < n_incl; jj++) dsum += dist[(size_t)ii * n_incl + jj];
      for (int jj = 0; jj < n_incl; jj++) dist[(size_t)ii * n_incl + jj] /= dsum;
    }
  }
}

/* relbase; src/Relbase.cpp:524-538 with relbase_profile :461-512 (no caching) */
int orc_relbase(const orc_reltable *tab, const orc_lptable *lp, const orc_relparam *p,
                const double *ener, int n_ener, int nr_fine, double *flux, double *dist,
                int n_incl, orc_stats *st) {
  int status = 0;
  orc_syspar *s = orc_get_system_parameters(tab, lp, p, nr_fine, &status);
  if (!s) return status;
  double *rgrid = (double *)malloc(sizeof(double) * (p->num_zones + 1));
  orc_get_rzone_grid(p->rin, p->rout, p->num_zones, p->height, rgrid);
  status = orc_calc_relline_profile(s, ener, n_ener, rgrid, p->num_zones, flux, dist, n_incl, st);
  if (status == 0) {
    orc_renorm_relline_profile(ener, n_ener, p->num_zones, flux, dist, n_incl, p);
  }
  free(rgrid);
  orc_syspar_free(s);
  return status;
}

/* ------------------------------------------------------------------------ */
/* FFT convolution (src/Relbase.cpp:93-246)                                  */
/* ------------------------------------------------------------------------ */

/* iterative radix-2 complex FFT, unnormalised: forward uses exp(-2 pi i jk/n),
 * inverse exp(+2 pi i jk/n); n must be a power of two */
void orc_fft(double *re, double *im, int n, int inverse) {
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len >> 1;
    for (int k = 0; k < half; k++) {
      double ang = 2.0 * M_PI * k / len * (inverse ? 1.0 : -1.0);
      double wr = cos(ang), wi = sin(ang);
      for (int i = k; i < n; i += len) {
        int j = i + half;
        double xr = re[j] * wr - im[j] * wi;
        double xi = re[j] * wi + im[j] * wr;
        re[j] = re[i] - xr;
        im[j] = im[i] - xi;
        re[i] += xr;
        im[i] += xi;
      }
    }
  }
}

/* fftw_conv_spectrum; src/Relbase.cpp:119-188 (with the E/dE conversion :93-103
 * and the 1 keV rotation :133-137,:161) */
void orc_fft_conv_spectrum(const double *ener, const double *fxill, const double *frel,
                           double *fout, int n) {
  double *c = (double *)malloc(sizeof(double) * n);
  double *xr = (double *)calloc(n, sizeof(double)), *xi = (double *)calloc(n, sizeof(double));
  double *yr = (double *)calloc(n, sizeof(double)), *yi = (double *)calloc(n, sizeof(double));
  for (int ii = 0; ii < n; ii++) c[ii] = 0.5 * (ener[ii] + ener[ii + 1]) / (ener[ii + 1] - ener[ii]);
  int pos1 = orc_binary_search(ener, n + 1, 1.0);
  for (int ii = 0; ii < n; ii++) {
    xr[ii] = fxill[ii] * c[ii];
    int irot = (ii - pos1 + n) % n;
    yr[irot] = frel[ii] * c[ii];
  }
  orc_fft(xr, xi, n, 0);
  orc_fft(yr, yi, n, 0);
  /* product on the n/2+1 Hermitian half, as c2r reads it */
  const int nh = n / 2;
  double *pr = (double *)calloc(n, sizeof(double)), *pi = (double *)calloc(n, sizeof(double));
  for (int ii = 0; ii <= nh; ii++) {
    pr[ii] = xr[ii] * yr[ii] - xi[ii] * yi[ii];
    pi[ii] = xr[ii] * yi[ii] + xi[ii] * yr[ii];
  }
  pi[0] = 0.0;
  pi[nh] = 0.0;
  for (int ii = 1; ii < nh; ii++) {
    pr[n - ii] = pr[ii];
    pi[n - ii] = -pi[ii];
  }
  orc_fft(pr, pi, n, 1);
  for (int ii = 0; ii < n; ii++) fout[ii] = pr[ii] / c[ii];
  free(c); free(xr); free(xi); free(yr); free(yi); free(pr); free(pi);
}

/* calcFFTNormFactor; src/Relbase.cpp:199-213 */
double orc_fft_norm_factor(const double *ener, const double *fxill, const double *frel,
                           const double *fout, int n) {
  double sum_relline = 0.0, sum_xillver = 0.0, sum_conv = 0.0;
  for (int jj = 0; jj < n; jj++) {
    if (ener[jj] >= EMIN_BAND && ener[jj + 1] < EMAX_BAND) {
      sum_xillver += fxill[jj];
      sum_relline += frel[jj];
      sum_conv += fout[jj];
    }
  }
  return sum_relline * sum_xillver / sum_conv;
}

/* convolveSpectrumFFTNormalized; src/Relbase.cpp:215-230 */
void orc_conv_normalized(const double *ener, const double *fxill, const double *frel,
                         double *fout, int n) {
  orc_fft_conv_spectrum(ener, fxill, frel, fout, n);
  double f = orc_fft_norm_factor(ener, fxill, frel, fout, n);
  for (int ii = 0; ii < n; ii++) fout[ii] *= f;
}

/* set_flux_outside_defined_range_to_zero; src/Relbase.cpp:233-246 */
void orc_zero_outside(const double *ener, double *spec, int n, double emin, double emax) {
  for (int ii = 0; ii < n; ii++) {
    if (ener[ii + 1] < emin || ener[ii] > emax) spec[ii] = 0;
  }
}

/* relconv_kernel; src/Relbase.cpp:257-289 */
int orc_relconv_kernel(const orc_reltable *tab, const orc_lptable *lp, const orc_relparam *p,
                       const double *conv_ener, int n_conv, const double *ener_inp,
                       double *spec_inp, int n_inp, int nr_fine, orc_stats *st) {
  double *rel = (double *)malloc(sizeof(double) * n_conv);
  double *rebin_flux = (double *)malloc(sizeof(double) * n_conv);
  double *conv_out = (double *)malloc(sizeof(double) * n_conv);
  int status = orc_relbase(tab, lp, p, conv_ener, n_conv, nr_fine, rel, NULL, 0, st);
  if (status == 0) {
    orc_rebin_spectrum(conv_ener, rebin_flux, n_conv, ener_inp, spec_inp, n_inp);
    orc_conv_normalized(conv_ener, rebin_flux, rel, conv_out, n_conv);
    orc_rebin_spectrum(ener_inp, spec_inp, n_inp, conv_ener, conv_out, n_conv);
    orc_zero_outside(ener_inp, spec_inp, n_inp, RELCONV_EMIN, RELCONV_EMAX);
  }
  free(rel); free(rebin_flux); free(conv_out);
  return status;
}

/* relxill-style, one zone: profile with angle distribution (relbase_profile
 * with a xillver grid, src/Relbase.cpp:461-512), angle weighting
 * (calc_xillver_angdep, src/Xillspec.cpp:557-573), per-zone rebin + FFT
 * convolution + accumulation (src/Relxill.cpp:432-482), rebin to the caller
 * grid (src/Relxill.cpp:261-278) */
int orc_relxill_conv(const orc_reltable *tab, const orc_relparam *p, const double *conv_ener,
                     int n_conv, const double *refl_ener, const double *refl, int n_refl,
                     int n_incl, const double *out_ener, double *out, int n_out, int nr_fine,
                     double *dist_out, orc_stats *st) {
  double *rel = (double *)malloc(sizeof(double) * n_conv);
  double *dist = (double *)malloc(sizeof(double) * n_incl);
  double *xill = (double *)malloc(sizeof(double) * n_refl);
  double *xill_rb = (double *)malloc(sizeof(double) * n_conv);
  double *conv_out = (double *)malloc(sizeof(double) * n_conv);
  double *total = (double *)calloc(n_conv, sizeof(double));
  int status = orc_relbase(tab, NULL, p, conv_ener, n_conv, nr_fine, rel, dist, n_incl, st);
  if (status == 0) {
    for (int jj = 0; jj < n_refl; jj++) xill[jj] = 0.0;
    for (int ii = 0; ii < n_incl; ii++) {
      for (int jj = 0; jj < n_refl; jj++) xill[jj] += dist[ii] * refl[(size_t)ii * n_refl + jj];
    }
    double relsum = 0.0;
    for (int jj = 0; jj < n_conv; jj++) relsum += rel[jj];
    if (!(relsum < 1e-12)) {
      orc_rebin_spectrum(conv_ener, xill_rb, n_conv, refl_ener, xill, n_refl);
      orc_conv_normalized(conv_ener, xill_rb, rel, conv_out, n_conv);
      for (int jj = 0; jj < n_conv; jj++) total[jj] += conv_out[jj];
    }
    orc_rebin_spectrum(out_ener, out, n_out, conv_ener, total, n_conv);
    if (dist_out) memcpy(dist_out, dist, sizeof(double) * n_incl);
  }
  free(rel); free(dist); free(xill); free(xill_rb); free(conv_out); free(total);
  return status;
}

/* ------------------------------------------------------------------------ */
/* xillver table and the relxill model                                       */
/* ------------------------------------------------------------------------ */

static const double EMIN_XILL_NORM = 0.1;     /* src/Xillspec.h:28 */
static const double EMAX_XILL_NORM = 1000.0;  /* src/Xillspec.h:29 */
static const int N_ENER_COARSE = 500;         /* src/Xillspec.h:32 */

/* the input value of each table parameter, as float (get_xilltab_paramvals,
 * src/xilltable.c:278-295; kTbb is 0 but for the blackbody model, Frac is 0) */
static void xill_paramvals(const orc_xillparam *x, float *v) {
  for (int k = 0; k < 8; k++) v[k] = 0.0f;
  v[ORC_XP_GAM] = (float)x->gam;
  v[ORC_XP_AFE] = (float)x->afe;
  v[ORC_XP_LXI] = (float)x->lxi;
  v[ORC_XP_ECT] = (float)x->ect;
  v[ORC_XP_DNS] = (float)x->dens;
  v[ORC_XP_KTB] = (float)x->ktbb;
  v[ORC_XP_INC] = (float)x->incl;
}

void orc_xill_indices(const orc_xilltable *t, const orc_xillparam *x, int *ind) {
  float v[8];
  xill_paramvals(x, v);
  for (int k = 0; k < t->n_par; k++) {
    ind[k] = orc_binary_search_float(t->val[k], t->n_val[k], v[t->pid[k]]);
    if (ind[k] < 0) ind[k] = 0;
    else if (ind[k] > t->n_val[k] - 2) ind[k] = t->n_val[k] - 2;
  }
}

static int xill_dim_of(const orc_xilltable *t, int id) {  /* get_xilltab_param_index, :455-465 */
  for (int k = 0; k < t->n_par; k++)
    if (t->pid[k] == id) return k;
  return -1;
}

void orc_xill_factors(const orc_xilltable *t, const orc_xillparam *x, const int *ind, double *fac) {
  float v[8];
  xill_paramvals(x, v);
  for (int k = 0; k < t->n_par; k++) {
    const int pk = t->pid[k];
    /* relxill models do not clamp the inclination (it is not interpolated) */
    if (pk != ORC_XP_INC) {
      const float lo = t->val[k][0], hi = t->val[k][t->n_val[k] - 1];
      if (v[pk] < lo) v[pk] = lo;
      else if (v[pk] > hi) v[pk] = hi;
    }
    /* float arithmetic throughout, as in the reference */
    const float num = v[pk] - t->val[k][ind[k]];
    const float den = t->val[k][ind[k] + 1] - t->val[k][ind[k]];
    fac[k] = (float)(num / den);
  }
  /* ensure_ecut_within_boundarys: the reference reads the Ecut axis as
   * param_vals[param_index[k_ect]], i.e. the axis whose number is the id of
   * Ecut (3); for tables with Ecut as fourth parameter that is its own axis */
  const int k_ect = xill_dim_of(t, ORC_XP_ECT);
  if (k_ect >= 0) {
    const int q = t->pid[k_ect];
    if (q < t->n_par) {
      if (x->ect <= t->val[q][0]) fac[k_ect] = 0.0;
      if (x->ect >= t->val[q][t->n_val[q] - 1]) fac[k_ect] = 1.0;
    }
  }
}

void orc_xill_renorm(float *spec, int n, double lxi, double dens) {
  for (int ii = 0; ii < n; ii++) {
    spec[ii] /= pow(10, lxi);
    if (fabs(dens - 15) > 1e-6) spec[ii] /= pow(10, dens - 15);
  }
}

/* one spectrum of the table, renormalised as when it is loaded
 * (xilltable_fits_load_single_spec, src/xilltable.c:513-566, with
 * normalizeXillverSpecLogxiDensity, :488-511) */
static void xill_row(const orc_xilltable *t, const orc_xillparam *x, const int *idx, float *dst) {
  size_t row = 0;
  for (int k = 0; k < t->n_par; k++) row = row * (size_t)t->n_val[k] + (size_t)idx[k];
  memcpy(dst, t->spec + row * (size_t)t->n_ener, sizeof(float) * t->n_ener);
  double dens = x->dens, lxi = x->lxi;
  const int k_dns = xill_dim_of(t, ORC_XP_DNS), k_lxi = xill_dim_of(t, ORC_XP_LXI);
  if (k_dns >= 0) dens = t->val[k_dns][idx[k_dns]];
  if (k_lxi >= 0) lxi = t->val[k_lxi][idx[k_lxi]];
  orc_xill_renorm(dst, t->n_ener, lxi, dens);
}

/* the corners of the four interpolated parameters in the reference's order
 * (interp_5d_tab_incl, src/xilltable.c:812-876; interp_5d_tab, :878-995) */
static const char corner[16][4] = {
    {0, 0, 0, 0}, {1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {1, 1, 0, 0}, {1, 0, 1, 0},
    {0, 1, 1, 0}, {1, 1, 1, 0}, {0, 0, 0, 1}, {1, 0, 0, 1}, {0, 1, 0, 1}, {0, 0, 1, 1},
    {1, 1, 0, 1}, {1, 0, 1, 1}, {0, 1, 1, 1}, {1, 1, 1, 1}};

/* interp_5d_tab_incl (src/xilltable.c:812-876) over the four axes after the
 * first `off` ones (which sit at idx0[]), for inclination ii: corners in the
 * reference's order, weights as its products */
static void xill_interp4(const orc_xilltable *t, const orc_xillparam *x, int off, const int *idx0, const int *ind,
                         const double *fac, int ii, float *rows, double *out) {
  const int n = t->n_ener;
  double w[16];
  for (int c = 0; c < 16; c++) {
    double f[4];
    for (int k = 0; k < 4; k++) f[k] = corner[c][k] ? fac[off + k] : (1.0 - fac[off + k]);
    w[c] = f[0] * f[1] * f[2] * f[3];
  }
  for (int c = 0; c < 16; c++) {
    int idx[6];
    for (int k = 0; k < off; k++) idx[k] = idx0[k];
    for (int k = 0; k < 4; k++) idx[off + k] = ind[off + k] + corner[c][k];
    idx[off + 4] = ii;
    xill_row(t, x, idx, rows + (size_t)c * n);
  }
  for (int e = 0; e < n; e++) {
    double v = 0.0;
    for (int c = 0; c < 16; c++) v = v + w[c] * (double)rows[(size_t)c * n + e];
    out[e] = v;
  }
}

void orc_xill_spectra(const orc_xilltable *t, const orc_xillparam *x, double *ener, double *flu) {
  const int n = t->n_ener, n_incl = t->n_val[t->n_par - 1];
  for (int ii = 0; ii < n; ii++) ener[ii] = t->elo[ii];
  ener[n] = t->ehi[n - 1];
  int ind[6];
  double fac[6];
  orc_xill_indices(t, x, ind);
  orc_xill_factors(t, x, ind, fac);
  float *rows = (float *)malloc(sizeof(float) * 16 * (size_t)n);
  if (t->n_par == 5) {
    for (int ii = 0; ii < n_incl; ii++) xill_interp4(t, x, 0, NULL, ind, fac, ii, rows, flu + (size_t)ii * n);
  } else {
    /* interp_6d_tab_incl (src/xilltable.c:997-1020): the five-parameter
     * interpolation at both brackets of the first axis, then
     * interp_lin_1d(fac[0], s1, s2) */
    double *s1 = (double *)malloc(sizeof(double) * 2 * (size_t)n), *s2 = s1 + n;
    for (int ii = 0; ii < n_incl; ii++) {
      int i0 = ind[0];
      xill_interp4(t, x, 1, &i0, ind, fac, ii, rows, s1);
      i0 = ind[0] + 1;
      xill_interp4(t, x, 1, &i0, ind, fac, ii, rows, s2);
      double *out = flu + (size_t)ii * n;
      for (int e = 0; e < n; e++) out[e] = fac[0] * s2[e] + (1.0 - fac[0]) * s1[e];
    }
    free(s1);
  }
  free(rows);
}

/* interp_5d_tab (:878-995) over the axes after the first `off` ones (at
 * idx0[]): 32 corners, the inclination's the fifth factor of each weight
 * product, summed in the reference's order */
static void xill_interp5(const orc_xilltable *t, const orc_xillparam *x, int off, const int *idx0, const int *ind,
                         const double *fac, float *rows, double *out) {
  const int n = t->n_ener, ki = t->n_par - 1;
  double w[32];
  for (int b5 = 0; b5 < 2; b5++) {
    for (int c = 0; c < 16; c++) {
      double f[4];
      for (int k = 0; k < 4; k++) f[k] = corner[c][k] ? fac[off + k] : (1.0 - fac[off + k]);
      w[16 * b5 + c] = f[0] * f[1] * f[2] * f[3] * (b5 ? fac[ki] : (1 - fac[ki]));
      int idx[6];
      for (int k = 0; k < off; k++) idx[k] = idx0[k];
      for (int k = 0; k < 4; k++) idx[off + k] = ind[off + k] + corner[c][k];
      idx[ki] = ind[ki] + b5;
      xill_row(t, x, idx, rows + (size_t)(16 * b5 + c) * n);
    }
  }
  for (int e = 0; e < n; e++) {
    double v = 0.0;
    for (int c = 0; c < 32; c++) v = v + w[c] * (double)rows[(size_t)c * n + e];
    out[e] = v;
  }
}

void orc_xill_spectrum_incl(const orc_xilltable *t, const orc_xillparam *x, double *ener, double *flu) {
  const int n = t->n_ener, ki = t->n_par - 1;
  for (int ii = 0; ii < n; ii++) ener[ii] = t->elo[ii];
  ener[n] = t->ehi[n - 1];
  int ind[6];
  double fac[6];
  orc_xill_indices(t, x, ind);
  orc_xill_factors(t, x, ind, fac);
  /* the xillver models clamp the inclination into the table as well
   * (resetInpvalsToBoundaries, src/xilltable.c:1117-1121); float arithmetic */
  float vi = (float)x->incl;
  const float ilo = t->val[ki][0], ihi = t->val[ki][t->n_val[ki] - 1];
  if (vi < ilo) vi = ilo;
  else if (vi > ihi) vi = ihi;
  {
    const float num = vi - t->val[ki][ind[ki]];
    const float den = t->val[ki][ind[ki] + 1] - t->val[ki][ind[ki]];
    fac[ki] = (float)(num / den);
  }
  float *rows = (float *)malloc(sizeof(float) * 32 * (size_t)n);
  if (t->n_par == 5) {
    xill_interp5(t, x, 0, NULL, ind, fac, rows, flu);
  } else {
    /* interp_6d_tab (src/xilltable.c:1022-1045) */
    double *s1 = (double *)malloc(sizeof(double) * 2 * (size_t)n), *s2 = s1 + n;
    int i0 = ind[0];
    xill_interp5(t, x, 1, &i0, ind, fac, rows, s1);
    i0 = ind[0] + 1;
    xill_interp5(t, x, 1, &i0, ind, fac, rows, s2);
    for (int e = 0; e < n; e++) flu[e] = fac[0] * s2[e] + (1.0 - fac[0]) * s1[e];
    free(s1);
  }
  free(rows);
}

/* LocalModel::xillver_model (src/LocalModel.cpp:104-128) with
 * shift_energy_grid_redshift (src/XspecSpectrum.h:70-76), norm_xillver_spec
 * (src/Xillspec.cpp:536-545) and add_primary_component (src/Relbase.cpp:
 * 294-344): the table spectrum at the parameters times 0.5 cos(Incl),
 * rebinned onto the redshifted grid, times |refl_frac|, plus the normalised
 * primary evaluated on that grid for refl_frac >= 0 */
int orc_xillver_model(int model_type, const orc_xilltable *xt, const double *par, const double *energy,
                      double *flux, int n_flux) {
  orc_xillparam x;
  memset(&x, 0, sizeof(x));
  double z;
  x.dens = 15;
  x.ect = 300;
  switch (model_type) {
    case ORC_MOD_XILLVER:  /* gamma Afe Ecut logxi z Incl refl_frac */
      x.gam = par[0]; x.afe = par[1]; x.ect = par[2]; x.lxi = par[3]; z = par[4]; x.incl = par[5];
      x.refl_frac = par[6];
      break;
    case ORC_MOD_XILLVERCP:  /* gamma Afe kTe logxi logN z Incl refl_frac */
      x.gam = par[0]; x.afe = par[1]; x.ect = par[2]; x.lxi = par[3]; x.dens = par[4]; z = par[5];
      x.incl = par[6]; x.refl_frac = par[7];
      x.prim = 1;
      break;
    case ORC_MOD_XILLVERNS:  /* kTbb Afe logN logxi z Incl refl_frac */
      x.ktbb = par[0]; x.afe = par[1]; x.dens = par[2]; x.lxi = par[3]; z = par[4]; x.incl = par[5];
      x.refl_frac = par[6];
      x.prim = 2;
      break;
    default:
      return 51;
  }
  const int nx = xt->n_ener;
  double *ener = (double *)malloc(sizeof(double) * (n_flux + 1));
  double *xe = (double *)malloc(sizeof(double) * (nx + 1));
  double *spec = (double *)malloc(sizeof(double) * nx);
  double *prim = (double *)malloc(sizeof(double) * n_flux);
  memcpy(ener, energy, sizeof(double) * (n_flux + 1));
  if (z > 0)
    for (int ii = 0; ii <= n_flux; ii++) ener[ii] *= (1 + z);
  orc_xill_spectrum_incl(xt, &x, xe, spec);
  const double slab = 0.5 * cos(x.incl * M_PI / 180);
  for (int ii = 0; ii < nx; ii++) spec[ii] *= slab;
  orc_rebin_spectrum(ener, flux, n_flux, xe, spec, nx);
  const double nf = orc_prim_norm_factor(&x, x.ect);
  orc_prim_spectrum(prim, ener, n_flux, &x, x.ect, 1.0);
  for (int ii = 0; ii < n_flux; ii++) prim[ii] *= nf;
  for (int ii = 0; ii < n_flux; ii++) flux[ii] *= fabs(x.refl_frac);
  if (x.refl_frac >= 0)
    for (int ii = 0; ii < n_flux; ii++) flux[ii] += prim[ii];
  free(ener); free(xe); free(spec); free(prim);
  return 0;
}

void orc_xill_angdep(double *out, const double *flu, const double *dist, int n_incl, int n_ener) {
  for (int jj = 0; jj < n_ener; jj++) out[jj] = 0.0;
  for (int ii = 0; ii < n_incl; ii++)
    for (int jj = 0; jj < n_ener; jj++) out[jj] += dist[ii] * flu[(size_t)ii * n_ener + jj];
}

/* corrected_gshift_fluxboost_factor; src/Relreturn_Corona.cpp:39-83 */
double orc_rrad_gshift_boost(double c, double g, double gamma) {
  const double g0 = 2. / 3;
  double corr;
  if (c < 1) { /* parabola through f(0) = 0 */
    const double a = (c / g0 - 1) / (g0 - 1);
    const double b = 1 - a;
    corr = (g >= 1) ? 1. / g * (1. / g * a + b) : g * (g * a + b);
  } else { /* linear */
    const double alin = (c - 1) / (g0 - 1);
    const double blin = 1 - alin;
    corr = (g >= 1) ? (1. / g * alin + blin) : (g * alin + blin);
  }
  double boost = pow(g, gamma) * corr;
  if (g < 1 && boost > 1) boost = 1;
  if (boost < 0) boost = 0;
  return boost;
}

/* get_energy_flux_band / get_photon_flux_band; src/Xillspec.cpp:132-162 */
static double energy_flux_band(const double *ener, const double *f, int n, double emin, double emax) {
  double sum = 0.0;
  for (int ii = 0; ii < n; ii++)
    if (ener[ii] >= emin && ener[ii + 1] <= emax) sum += f[ii] * 0.5 * (ener[ii] + ener[ii + 1]);
  return sum;
}

static double photon_flux_band(const double *ener, const double *f, int n, double emin, double emax) {
  double sum = 0.0;
  for (int ii = 0; ii < n; ii++)
    if (ener[ii] >= emin && ener[ii + 1] <= emax) sum += f[ii];
  return sum;
}

/* get_xillver_fluxcorrection_factors; src/Xillspec.cpp:500-526, with
 * calc_angle_averaged_xill_spec (:109-126), get_xillver_fluxcorr (:344-362)
 * and get_xillver_gshift_fluxcorr (:454-487) */
void orc_xill_corr_factors(const orc_xilltable *t, const orc_xillparam *x, double *corr_flux, double *corr_gshift) {
  const int nx = t->n_ener, n_incl = t->n_val[t->n_par - 1];
  const float *incl = t->val[t->n_par - 1];
  double *e = (double *)malloc(sizeof(double) * (nx + 1));
  double *flu = (double *)malloc(sizeof(double) * (size_t)n_incl * nx);
  double *avg = (double *)calloc(nx, sizeof(double));
  orc_xill_spectra(t, x, e, flu);
  /* the angle average: the weight of an inclination is
   * norm_factor_semi_infinite_slab of the table's value (in degrees, as
   * the xillver spectrum keeps it, src/xilltable.c:1113) times 180/pi,
   * i.e. 0.5 cos of the degree value read as radians, over n_incl */
  for (int ii = 0; ii < n_incl; ii++) {
    const double incl_deg = (double)incl[ii] * 180 / M_PI;
    const double w = 0.5 * cos(incl_deg * M_PI / 180) / n_incl;
    for (int jj = 0; jj < nx; jj++) avg[jj] += w * flu[(size_t)ii * nx + jj];
  }
  /* reflected over direct energy flux in 0.1-1000 keV, the direct spectrum
   * the normalised cut-off power law on the coarse grid */
  double coarse[N_ENER_COARSE + 1], prim[N_ENER_COARSE];
  orc_get_log_grid(coarse, N_ENER_COARSE + 1, EMIN_XILL_NORM, EMAX_XILL_NORM);
  orc_prim_spectrum(prim, coarse, N_ENER_COARSE, x, x->ect, 1.0);
  int st = 0;
  const double nf = 1. / orc_norm_wrt_xillver(prim, coarse, N_ENER_COARSE, &st);
  for (int ii = 0; ii < N_ENER_COARSE; ii++) prim[ii] *= nf;
  *corr_flux = energy_flux_band(e, avg, nx, 0.1, 1000) / energy_flux_band(coarse, prim, N_ENER_COARSE, 0.1, 1000);
  /* photon flux of the spectrum over that of the spectrum shifted to 1.5
   * times lower energies (rebinned, times 2/3 for the time dilation), in
   * 0.15-500 keV, over the power law's 1.5^Gamma */
  const double gref = 2. / 3.;
  double *ez = (double *)malloc(sizeof(double) * (nx + 1));
  double *fz = (double *)malloc(sizeof(double) * nx);
  for (int ii = 0; ii <= nx; ii++) ez[ii] = e[ii] / gref;
  orc_rebin_spectrum(ez, fz, nx, e, avg, nx);
  for (int ii = 0; ii < nx; ii++) fz[ii] *= gref;
  const double ratio = photon_flux_band(e, avg, nx, 0.15, 500.0) / photon_flux_band(e, fz, nx, 0.15, 500.0);
  *corr_gshift = ratio / pow(1.5, x->gam);
  free(e); free(flu); free(avg); free(ez); free(fz);
}

void orc_spec_cutoffpl(double *flux, const double *ener, int n, double ect, double gam, double shift) {
  const double ecut = ect * shift;
  for (int ii = 0; ii < n; ii++) {
    const double en = 0.5 * (ener[ii] + ener[ii + 1]);
    flux[ii] = exp(1.0 / ecut) * pow(en, -gam) * exp(-en / ecut) * (ener[ii + 1] - ener[ii]);
  }
}

double orc_norm_wrt_xillver(const double *flux, const double *ener, int n, int *status) {
  const double keV2erg = 1.602177e-09;
  const double tol = 1e-6;
  if (ener[n] < (EMAX_XILL_NORM - tol) || ener[0] > (EMIN_XILL_NORM + tol)) {
    if (status) *status = 1;
    return 0.0;
  }
  double sum_pl = 0.0;
  for (int ii = 0; ii < n; ii++) {
    if (ener[ii] >= EMIN_XILL_NORM && ener[ii] <= EMAX_XILL_NORM)
      sum_pl += flux[ii] * 0.5 * (ener[ii] + ener[ii + 1]) * 1e20 * keV2erg;
  }
  const double norm_xillver_table = 1e15 / 4.0 / M_PI;
  return sum_pl / norm_xillver_table;
}

double orc_primary_norm_factor(double gam, double ect) {
  double ener[N_ENER_COARSE + 1], flux[N_ENER_COARSE];
  orc_get_log_grid(ener, N_ENER_COARSE + 1, EMIN_XILL_NORM, EMAX_XILL_NORM);
  orc_spec_cutoffpl(flux, ener, N_ENER_COARSE, ect, gam, 1.0);
  return 1. / orc_norm_wrt_xillver(flux, ener, N_ENER_COARSE, NULL);
}

void orc_spec_blackbody(double *flux, const double *ener, int n, double ktbb) {
  for (int ii = 0; ii < n; ii++) {
    const double en = 0.5 * (ener[ii] + ener[ii + 1]);
    flux[ii] = en * en / (pow(ktbb, 4) * (exp(en / ktbb) - 1));
    flux[ii] *= (ener[ii + 1] - ener[ii]);
  }
}

/* calc_primary_spectrum (src/Xillspec.cpp:278-300): the blackbody takes no
 * energy shift */
void orc_prim_spectrum(double *flux, const double *ener, int n, const orc_xillparam *x, double ect, double shift) {
  if (x->prim == 2) orc_spec_blackbody(flux, ener, n, x->ktbb);
  else if (x->prim == 1) orc_spec_nthcomp(flux, ener, n, ect, x->gam, shift);
  else orc_spec_cutoffpl(flux, ener, n, ect, x->gam, shift);
}

double orc_prim_norm_factor(const orc_xillparam *x, double ect) {
  double ener[N_ENER_COARSE + 1], flux[N_ENER_COARSE];
  orc_get_log_grid(ener, N_ENER_COARSE + 1, EMIN_XILL_NORM, EMAX_XILL_NORM);
  orc_prim_spectrum(flux, ener, N_ENER_COARSE, x, ect, 1.0);
  return 1. / orc_norm_wrt_xillver(flux, ener, N_ENER_COARSE, NULL);
}

/* relxillCp / relxilllpCp (lmodel_relxill_public.dat:96-129; convertModelType,
 * src/ModelDefinition.cpp:38-50): the parameters in relxill's / relxilllp's
 * order with kTe in Ecut's place; returns that model, *dens logN and
 * relxilllpCp's iongrad_type and iongrad_index */
static int cp_remap(int model_type, const double *par, double *out, double *dens, int *iongrad,
                    double *ion_index) {
  /* relxillCp: Incl a Rin Rout Rbr Index1 Index2 z gamma logxi logN Afe kTe refl_frac */
  static const int k_bkn[13] = {5, 6, 4, 1, 0, 2, 3, 7, 8, 9, 11, 12, 13};
  /* relxilllpCp: Incl a Rin Rout h beta gamma logxi logN Afe kTe refl_frac z
   * iongrad_index iongrad_type switch_returnrad switch_reflfrac_boost */
  static const int k_lp[14] = {4, 5, 1, 0, 2, 3, 12, 6, 7, 9, 10, 11, 15, 16};
  *iongrad = 0;
  if (model_type == ORC_MOD_RELXILLCP) {
    for (int i = 0; i < 13; i++) out[i] = par[k_bkn[i]];
    *dens = par[10];
    return ORC_MOD_RELXILL;
  }
  for (int i = 0; i < 14; i++) out[i] = par[k_lp[i]];
  *dens = par[8];
  *iongrad = (int)lround(par[14]);
  *ion_index = par[13];
  return ORC_MOD_RELXILLLP;
}

int orc_get_relxill_params(int model_type, const double *par, int phys_norm_env, orc_relparam *p,
                           orc_xillparam *x) {
  double cpar[14], dens = 15, ktbb = 0, ion_index = 0;
  int prim = 0, iongrad = 0;
  if (model_type == ORC_MOD_RELXILLCP || model_type == ORC_MOD_RELXILLLPCP) {
    model_type = cp_remap(model_type, par, cpar, &dens, &iongrad, &ion_index);
    par = cpar;
    prim = 1;
  } else if (model_type == ORC_MOD_RELXILLNS) {
    /* relxillNS: Index1 Index2 Rbr a Incl Rin Rout z kTbb logxi Afe logN
     * refl_frac; no Gamma (0) and no Ecut (300, its default) in get_xill_params
     * (src/ModelDefinition.cpp:357-376) */
    for (int i = 0; i < 13; i++) cpar[i] = par[i];
    ktbb = par[8];
    dens = par[11];
    cpar[8] = 0.0;
    cpar[11] = 300.0;
    par = cpar;
    model_type = ORC_MOD_RELXILL;
    prim = 2;
  }
  int status = orc_get_rel_params(model_type, par, phys_norm_env, p);
  /* get_xill_params (src/ModelDefinition.cpp:339-390): Ecut in keV, the
   * density at its default 15; logXi kept within 0-4.7 for the zones
   * (adjust_parameter_to_be_within_xillver_bounds, src/IonGradient.cpp:36-44,
   * :388) */
  memset(x, 0, sizeof(*x));
  if (model_type == ORC_MOD_RELXILLLP) {
    x->gam = par[7]; x->lxi = par[8]; x->afe = par[9]; x->ect = par[10]; x->refl_frac = par[11];
    x->incl = par[3];
    x->boost = (int)lround(par[13]);
  } else {
    x->gam = par[8]; x->lxi = par[9]; x->afe = par[10]; x->ect = par[11]; x->refl_frac = par[12];
    x->incl = par[4];
  }
  /* logN: 15 for the cut-off power-law models, the parameter for the nthcomp ones
   * (get_xill_params, src/ModelDefinition.cpp:357-368) */
  x->dens = dens;
  x->prim = prim;
  x->ktbb = ktbb;
  p->ion_grad_type = 0;
  p->ion_grad_index = 0.0;
  p->lxi0 = x->lxi;
  if (x->lxi < 0.0) x->lxi = 0.0;
  else if (x->lxi > 4.7) x->lxi = 4.7;
  /* relxilllpCp's ionisation gradient (IonGradient::calculate_gradient,
   * src/IonGradient.cpp:340-395): the power law (1) on N_ZONES_IONGRAD zones
   * (get_num_zones, src/relutility.c:517-527); the alpha-disc gradient (2)
   * is not restated */
  if (status == 0 && iongrad == 1) {
    p->ion_grad_type = 1;
    p->ion_grad_index = ion_index;
    p->num_zones = 25;
  } else if (status == 0 && iongrad != 0) {
    status = 2;
  }
  return status;
}

/* energy_shift_source_obs (src/Relphysics.cpp:239-254) with calc_g_inf
 * (:134-136) and doppler_factor_source_obs (:217-230) */
double orc_energy_shift_source_obs(const orc_relparam *p) {
  if (p->emis_type != ORC_EMIS_LP) return 1;
  const double g_inf_0 = sqrt(1.0 - (2 * p->height / (p->height * p->height + p->a * p->a)));
  if (p->beta < 1e-4) return g_inf_0;
  return g_inf_0 * doppler_factor(M_PI - p->incl, p->beta);
}

void orc_zone_lxi(const orc_relparam *p, double lxi, const double *rmean, int nz, double *out) {
  for (int iz = 0; iz < nz; iz++) out[iz] = lxi;
  if (p->ion_grad_type != 1) return;
  /* calc_ion_grad_pl (src/IonGradient.cpp:174-182), then
   * adjust_parameter_to_be_within_xillver_bounds (:36-44, :388) */
  for (int iz = 0; iz < nz; iz++) {
    double v = exp(p->lxi0) * pow(rmean[iz] / rmean[0], -1.0 * p->ion_grad_index);
    v = log(v);
    if (v < 0.0) v = 0.0;
    else if (v > 4.7) v = 4.7;
    out[iz] = v;
  }
}

int orc_relxill_model(int model_type, const orc_reltable *tab, const orc_lptable *lp,
                      const orc_xilltable *xt, const double *parameter, const double *energy, double *flux,
                      int n_flux, const double *conv_ener, int n_conv, int phys_norm_env, double *dist_out,
                      orc_stats *st) {
  orc_relparam p;
  orc_xillparam x;
  int status = orc_get_relxill_params(model_type, parameter, phys_norm_env, &p, &x);
  if (status == 0 && p.return_rad != 0 && !(p.return_rad == 1 || p.return_rad == -1 || p.return_rad == 2))
    status = 34;
  if (status) {
    memset(flux, 0, sizeof(double) * n_flux);
    return status;
  }
  double *grid = NULL;
  if (!conv_ener) {  /* get_relxill_conv_energy_grid, src/Xillspec.cpp:44-54 */
    n_conv = 4096;
    grid = (double *)malloc(sizeof(double) * (n_conv + 1));
    orc_get_log_grid(grid, n_conv + 1, 0.00035, 2000.0);
    conv_ener = grid;
  }
  const int nz = p.num_zones;
  const int nx = xt->n_ener, n_incl = xt->n_val[xt->n_par - 1];
  double *ener = (double *)malloc(sizeof(double) * (n_flux + 1));
  double *xe = (double *)malloc(sizeof(double) * (nx + 1));
  double *xflu = (double *)malloc(sizeof(double) * (size_t)n_incl * nx);
  double *xill = (double *)malloc(sizeof(double) * nx);
  double *rel = (double *)malloc(sizeof(double) * (size_t)nz * n_conv);
  double *dist = (double *)calloc((size_t)nz * n_incl, sizeof(double));
  double *xill_rb = (double *)malloc(sizeof(double) * n_conv);
  double *conv_out = (double *)malloc(sizeof(double) * n_conv);
  double *total = (double *)calloc(n_conv, sizeof(double));
  double *prim = (double *)malloc(sizeof(double) * n_conv);
  double *rgrid = (double *)malloc(sizeof(double) * (nz + 1));
  double *rmean = (double *)malloc(sizeof(double) * nz);
  double *del_emit = (double *)malloc(sizeof(double) * nz);
  double *shift = (double *)malloc(sizeof(double) * nz);
  double *zlxi = (double *)malloc(sizeof(double) * nz);
  /* shift_energy_grid_redshift (src/XspecSpectrum.h:70-76) */
  memcpy(ener, energy, sizeof(double) * (n_flux + 1));
  if (p.z > 0)
    for (int ii = 0; ii <= n_flux; ii++) ener[ii] *= (1 + p.z);
  /* the lamp post gives Ecut in the observer frame: the source's is Ecut / g
   * (src/Relxill.cpp:197-200, src/PrimarySource.h:204-216); kTe of the
   * nthcomp models is the source's already */
  const double g_so = orc_energy_shift_source_obs(&p);
  orc_xillparam xs = x;
  if (p.emis_type == ORC_EMIS_LP && !x.prim) xs.ect /= g_so;
  orc_syspar *sp = orc_get_system_parameters(tab, lp, &p, 1000, &status);
  double *cf_flux = (double *)malloc(sizeof(double) * nz), *cf_g = (double *)malloc(sizeof(double) * nz);
  if (sp) {
    /* 1: zones (RadialGrid, src/IonGradient.cpp:457-489; IonGradient.h:63-70)
     * with their del_emit and source -> disk energy shift (:182-193, :421-430) */
    orc_get_rzone_grid(p.rin, p.rout, nz, p.height, rgrid);
    for (int iz = 0; iz < nz; iz++) rmean[iz] = 0.5 * (rgrid[iz] + rgrid[iz + 1]);
    for (int iz = 0; iz < nz; iz++) {
      shift[iz] = 1.0;
      del_emit[iz] = 0.0;
    }
    orc_zone_lxi(&p, x.lxi, rmean, nz, zlxi);
    if (p.emis_type == ORC_EMIS_LP) {
      inv_rebin_mean(sp->re, sp->del_emit, sp->nr, rmean, del_emit, nz);
      for (int iz = 0; iz < nz; iz++) shift[iz] = orc_gi_potential_lp(rmean[iz], p.a, p.height, p.beta, del_emit[iz]);
    }
    /* 3: returning radiation above spin 0 (SPIN_MIN_RRAD_CALC_CORRFAC,
     * src/Relxill.h:32): the xillver correction factors of every zone from
     * its reflection spectra (calc_rrad_corr_factors, src/Relxill.cpp:164-182,
     * :337-341), then the emissivity again with them (:344); the zones'
     * angles and shifts do not depend on the emissivity */
    if (p.return_rad != 0 && p.a > 0.0) {
      for (int iz = 0; iz < nz; iz++) {
        orc_xillparam xz = xs;
        xz.ect = xs.ect * shift[iz];
        xz.lxi = zlxi[iz];
        orc_xill_corr_factors(xt, &xz, &cf_flux[iz], &cf_g[iz]);
      }
      p.rr_corr_flux = cf_flux;
      p.rr_corr_gshift = cf_g;
      p.rr_corr_rgrid = rgrid;
      p.rr_corr_n = nz;
      orc_syspar_free(sp);
      sp = orc_get_system_parameters(tab, lp, &p, 1000, &status);
    }
  }
  if (sp) {
    /* 4: the relline profile of every zone with its angle distribution */
    status = orc_calc_relline_profile(sp, conv_ener, n_conv, rgrid, nz, rel, dist, n_incl, st);
    if (status == 0) orc_renorm_relline_profile(conv_ener, n_conv, nz, rel, dist, n_incl, &p);
  }
  if (status == 0) {
    /* normalisation change source -> disk (calc_xillver_normalization_change_source_to_disk,
     * src/Xillspec.cpp:408-435) */
    const double nf_src = orc_prim_norm_factor(&xs, xs.ect);
    for (int iz = 0; iz < nz; iz++) {
      /* 2 + 5: the zone's xillver spectra (Ecut seen by the zone,
       * get_xill_param_zone, src/IonGradient.cpp:196-209), angle weighting */
      orc_xillparam xz = xs;
      xz.ect = xs.ect * shift[iz];
      xz.lxi = zlxi[iz];
      orc_xill_spectra(xt, &xz, xe, xflu);
      orc_xill_angdep(xill, xflu, dist + (size_t)iz * n_incl, n_incl, nx);
      const double ncf = orc_prim_norm_factor(&xs, xs.ect * shift[iz]) / nf_src;
      for (int jj = 0; jj < nx; jj++) xill[jj] /= ncf;
      /* 6: relxill_convolution_multizone (src/Relxill.cpp:432-482) */
      const double *rz = rel + (size_t)iz * n_conv;
      double relsum = 0.0;
      for (int jj = 0; jj < n_conv; jj++) relsum += rz[jj];
      if (relsum < 1e-12) continue;
      orc_rebin_spectrum(conv_ener, xill_rb, n_conv, xe, xill, nx);
      orc_conv_normalized(conv_ener, xill_rb, rz, conv_out, n_conv);
      for (int jj = 0; jj < n_conv; jj++) total[jj] += conv_out[jj];
    }
    /* add_primary_spectrum (src/PrimarySource.cpp:66-125): the observed
     * primary (the cut-off power law with the source's Ecut shifted by g, or
     * nthcomp at redshift 1/g - 1), normalised as the xillver incident
     * spectrum at the source (calc_normalized_primary_spectrum,
     * src/Xillspec.cpp:302-332) */
    orc_prim_spectrum(prim, conv_ener, n_conv, &xs, xs.ect, g_so);
    for (int jj = 0; jj < n_conv; jj++) prim[jj] *= nf_src;
    if (p.emis_type != ORC_EMIS_LP) {
      for (int jj = 0; jj < n_conv; jj++) total[jj] *= fabs(x.refl_frac);
    } else {
      const orc_reflfrac *lr = &sp->refl;
      double refl_frac_input = x.refl_frac;
      if (x.boost) refl_frac_input *= lr->refl_frac;
      /* flux_boost_source_to_observer (src/PrimarySource.h:84-93) */
      double prime_fac = lr->f_inf_rest / 0.5 * pow(g_so, xs.gam);
      if (p.beta > 1e-4) prime_fac *= pow(doppler_factor(M_PI - p.incl, p.beta), 2);
      const double norm_fac_refl = (fabs(refl_frac_input)) / lr->refl_frac;
      for (int jj = 0; jj < n_conv; jj++) total[jj] *= norm_fac_refl;
      for (int jj = 0; jj < n_conv; jj++) prim[jj] *= prime_fac;
    }
    if (x.refl_frac >= 0)
      for (int jj = 0; jj < n_conv; jj++) total[jj] += prim[jj];
    /* rebin_and_normalize_relxill_for_xspec (src/Relxill.cpp:267-278) */
    orc_rebin_spectrum(ener, flux, n_flux, conv_ener, total, n_conv);
    if (dist_out) memcpy(dist_out, dist, sizeof(double) * (size_t)nz * n_incl);
  } else {
    memset(flux, 0, sizeof(double) * n_flux);
  }
  if (sp) orc_syspar_free(sp);
  free(cf_flux); free(cf_g);
  free(grid); free(ener); free(xe); free(xflu); free(xill); free(rel); free(dist);
  free(xill_rb); free(conv_out); free(total); free(prim); free(rgrid); free(rmean); free(del_emit);
  free(shift);
  free(zlxi);
  return status;
}

int orc_relxill_zones(int model_type, const orc_reltable *tab, const orc_lptable *lp, const double *parameter,
                      double *out) {
  orc_relparam p;
  orc_xillparam x;
  int status = orc_get_relxill_params(model_type, parameter, 0, &p, &x);
  if (status) return status;
  const int nz = p.num_zones;
  orc_syspar *sp = orc_get_system_parameters(tab, lp, &p, 1000, &status);
  if (!sp) return status;
  double rgrid[64], rmean[64], del[64];
  orc_get_rzone_grid(p.rin, p.rout, nz, p.height, rgrid);
  for (int iz = 0; iz < nz; iz++) {
    rmean[iz] = 0.5 * (rgrid[iz] + rgrid[iz + 1]);
    del[iz] = 0.0;
  }
  if (p.emis_type == ORC_EMIS_LP) inv_rebin_mean(sp->re, sp->del_emit, sp->nr, rmean, del, nz);
  const double g_so = orc_energy_shift_source_obs(&p);
  const double ect_src = (p.emis_type == ORC_EMIS_LP && !x.prim) ? x.ect / g_so : x.ect;
  const double nf_src = orc_prim_norm_factor(&x, ect_src);
  double zlxi[64];
  orc_zone_lxi(&p, x.lxi, rmean, nz, zlxi);
  for (int iz = 0; iz < nz; iz++) {
    const double shift =
        (p.emis_type == ORC_EMIS_LP) ? orc_gi_potential_lp(rmean[iz], p.a, p.height, p.beta, del[iz]) : 1.0;
    double *o = out + 7 * iz;
    o[0] = rgrid[iz];
    o[1] = rgrid[iz + 1];
    o[2] = del[iz];
    o[3] = shift;
    o[4] = ect_src * shift;
    o[5] = orc_prim_norm_factor(&x, ect_src * shift) / nf_src;
    o[6] = zlxi[iz];
  }
  orc_syspar_free(sp);
  return 0;
}

int orc_relxill(const orc_reltable *tab, const orc_xilltable *xt, const double *parameter,
                const double *energy, double *flux, int n_flux, const double *conv_ener,
                int n_conv, int phys_norm_env, double *dist_out, orc_stats *st) {
  return orc_relxill_model(ORC_MOD_RELXILL, tab, NULL, xt, parameter, energy, flux, n_flux, conv_ener, n_conv,
                           phys_norm_env, dist_out, st);
}

/* ------------------------------------------------------------------------ */
/* model level (src/LocalModel.cpp:32-98, src/LocalModel.h:108-128)          */
/* ------------------------------------------------------------------------ */

int orc_eval_model(int model_type, const orc_reltable *tab, const orc_lptable *lp,
                   const double *parameter, const double *energy, double *flux, int n_flux,
                   const double *conv_ener, int n_conv, int phys_norm_env, orc_stats *st) {
  orc_relparam p;
  int status = orc_get_rel_params(model_type, parameter, phys_norm_env, &p);
  if (status) return status;
  double *ener = (double *)malloc(sizeof(double) * (n_flux + 1));
  memcpy(ener, energy, sizeof(double) * (n_flux + 1));
  /* shift_energy_grid_redshift (src/XspecSpectrum.h:70-76) */
  if (p.z > 0) {
    for (int ii = 0; ii <= n_flux; ii++) ener[ii] *= (1 + p.z);
  }
  const int nr_fine = 1000;  /* N_FRAD, src/common.h:133 */
  if (model_type == ORC_MOD_RELLINE || model_type == ORC_MOD_RELLINELP) {
    for (int ii = 0; ii <= n_flux; ii++) ener[ii] /= p.lineE;  /* shift_energy_grid_1keV */
    status = orc_relbase(tab, lp, &p, ener, n_flux, nr_fine, flux, NULL, 0, st);
  } else if (model_type == ORC_MOD_RELCONV || model_type == ORC_MOD_RELCONVLP) {
    double sum = 0.0;
    for (int ii = 0; ii < n_flux; ii++) sum += flux[ii];
    if (sum <= 0.0) {
      status = 50;
    } else {
      double *grid = NULL;
      if (!conv_ener) {  /* get_relxill_conv_energy_grid, src/Xillspec.cpp:44-54 */
        n_conv = 4096;
        grid = (double *)malloc(sizeof(double) * (n_conv + 1));
        orc_get_log_grid(grid, n_conv + 1, 0.00035, 2000.0);
        conv_ener = grid;
      }
      status = orc_relconv_kernel(tab, lp, &p, conv_ener, n_conv, ener, flux, n_flux, nr_fine, st);
      free(grid);
    }
  } else {
    status = 51;
  }
  free(ener);
  if (status) memset(flux, 0, sizeof(double) * n_flux);
  return status;
}

/* ------------------------------------------------------------------------ */
/* multi-threaded batch (CPU baseline: one spectrum per thread at a time)    */
/* ------------------------------------------------------------------------ */

typedef struct {
  const orc_reltable *tab;
  const orc_lptable *lp;
  orc_job *jobs;
  int n_jobs;
  const double *conv_ener;
  int n_conv;
  int *next;
  pthread_mutex_t *mu;
  orc_stats *total; /* optional: work counters summed over the jobs */
} mt_ctx;

static void *mt_worker(void *arg) {
  mt_ctx *c = (mt_ctx *)arg;
  for (;;) {
    pthread_mutex_lock(c->mu);
    int i = (*c->next)++;
    pthread_mutex_unlock(c->mu);
    if (i >= c->n_jobs) break;
    orc_job *j = &c->jobs[i];
    orc_stats st;
    memset(&st, 0, sizeof(st));
    j->status = orc_eval_model(j->model_type, c->tab, c->lp, j->parameter, j->energy, j->flux,
                               j->n_flux, c->conv_ener, c->n_conv, 0, c->total ? &st : NULL);
    if (c->total) {
      pthread_mutex_lock(c->mu);
      c->total->relb_func_calls += st.relb_func_calls;
      c->total->romberg_calls += st.romberg_calls;
      c->total->romberg_levels += st.romberg_levels;
      c->total->edge_calls += st.edge_calls;
      c->total->midpoint_calls += st.midpoint_calls;
      c->total->bin_integrals += st.bin_integrals;
      pthread_mutex_unlock(c->mu);
    }
  }
  return NULL;
}

int orc_eval_batch_mt_stats(const orc_reltable *tab, const orc_lptable *lp, orc_job *jobs, int n_jobs,
                            const double *conv_ener, int n_conv, int n_threads, orc_stats *total) {
  if (n_threads < 1) n_threads = 1;
  int next = 0;
  pthread_mutex_t mu;
  pthread_mutex_init(&mu, NULL);
  if (total) memset(total, 0, sizeof(*total));
  mt_ctx ctx = {tab, lp, jobs, n_jobs, conv_ener, n_conv, &next, &mu, total};
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * n_threads);
  for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, mt_worker, &ctx);
  for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
  free(th);
  pthread_mutex_destroy(&mu);
  int bad = 0;
  for (int i = 0; i < n_jobs; i++) bad += (jobs[i].status != 0);
  return bad;
}

int orc_eval_batch_mt(const orc_reltable *tab, const orc_lptable *lp, orc_job *jobs, int n_jobs,
                      const double *conv_ener, int n_conv, int n_threads) {
  return orc_eval_batch_mt_stats(tab, lp, jobs, n_jobs, conv_ener, n_conv, n_threads, NULL);
}
