/*
 * nthcomp_oracle.c -- FP64 CPU oracle of the nthcomp primary continuum.
 *
 * TEST INFRASTRUCTURE ONLY (see relxill_oracle.c): the checker of the
 * library's nthcomp kernel and of the relxillCp model.
 *
 * Restates the thermal Comptonization continuum that relxill uses as the
 * primary spectrum of its "Cp" models (spec_nthcomp, src/Xillspec.cpp:243-253,
 * with get_nthcomp_param, src/relutility.c:625-632: seed kT_bb = 0.05 keV,
 * disc-blackbody seed, redshift from the energy shift) as computed by
 * c_donthcomp (src/donthcomp.c:683-793): the Kompaneets equation with
 * relativistic corrections (Lightman & Zdziarski 1987; Zdziarski, Johnson &
 * Magdziarz 1996) solved on a log photon-energy grid as one tridiagonal
 * system, the seed a multicolour disc (Mitsuda et al. 1984) from the
 * interpolation formula of the diskbb model.  The arithmetic follows the
 * reference's order of operations (it defines the result at 1e-8); arrays
 * are 0-based here, and c_donthcomp's prim[] has its ne + 1 entries (the
 * reference's stack array is one short).
 */
#include <math.h>
#include <string.h>

#include "relxill_oracle.h"

#define NTH_MAX 900

static const double LOG10E = 0.43429448190325182765;  /* d_lg10, src/donthcomp.c:41-45 */
static double lg10(double x) { return LOG10E * log(x); }

/* Multicolour-disc photon spectrum per unit Tin^2 at et = E / Tin, from the
 * diskbb interpolation formula (f_mcdint__, src/donthcomp.c:50-115): a fitted
 * power law times exp(-et), three Gaussians in log10(et) and a residual table
 * on a 0.06-dex grid from et = 1e-3.  The 98 residuals and the Gaussians'
 * centres, widths and amplitudes are the model's published coefficients. */
static const double MCD_GC[3] = {.078196667, -1.066202, 1.192418};
static const double MCD_GW[3] = {.5207874, .513457, .4077983};
static const double MCD_GN[3] = {.3728691, .039775528, .037766505};
static const double MCD_RES[98] = {
    9.6198382e-4,   .0010901181,    .0012310012,    .0013841352,    .0015481583,    .0017210036,
    .0018988943,    .002076939,     .0022484281,    .0024049483,    .0025366202,    .0026316255,
    .0026774985,    .0026613059,    .0025708784,    .0023962965,    .002130655,     .0017725174,
    .0013268656,    8.0657672e-4,   2.3337584e-4,   -3.6291778e-4,  -9.4443569e-4,  -.0014678875,
    -.0018873741,   -.0021588493,   -.0022448371,   -.0021198179,   -.0017754602,   -.0012246034,
    -5.0414167e-4,  3.2507078e-4,   .0011811065,    .0019673402,    .0025827094,    .0029342526,
    .0029517083,    .0026012166,    .0018959062,    9.0128649e-4,   -2.6757144e-4,  -.0014567885,
    -.002492855,    -.0032079776,   -.0034678637,   -.0031988217,   -.0024080969,   -.001193624,
    2.6134145e-4,   .0017117758,    .0028906898,    .0035614435,    .0035711778,    .0028921374,
    .0016385898,    4.9857464e-5,   -.0015572671,   -.0028578151,   -.0035924212,   -.0036253044,
    -.002975086,    -.0018044436,   -3.7796664e-4,  .0010076215,    .0020937327,    .0027090854,
    .0028031667,    .0024276576,    .0017175597,    8.1030795e-4,   -1.2592304e-4,  -9.4888491e-4,
    -.0015544816,   -.0018831972,   -.0019203142,   -.0016905849,   -.0012487737,   -6.6789911e-4,
    -2.7079461e-5,  5.9931935e-4,   .0011499748,    .0015816521,    .0018709224,    .0020129966,
    .0020184702,    .0019089181,    .0017122289,    .001458377,     .0011760717,    8.9046768e-4,
    6.2190822e-4,   3.8553762e-4,   1.9155022e-4,   4.5837109e-5,   -4.9177834e-5,  -9.3670762e-5,
    -8.9622968e-5,  -4.01538532e-5};

double orc_mcd_shape(double et) {
  const double loget = lg10(et);
  double pos = (loget - lg10(0.001)) / .06 + 1;
  const int j = (int)pos;
  double resfact;
  if (j < 1) {
    resfact = MCD_RES[0];
  } else if (j >= 98) {
    resfact = MCD_RES[97];
  } else {
    pos -= j;
    resfact = MCD_RES[j - 1] * (1. - pos) + MCD_RES[j] * pos;
  }
  double gaufact = 1.;
  for (int k = 0; k < 3; k++) {
    const double z = (loget - MCD_GC[k]) / MCD_GW[k];
    gaufact += MCD_GN[k] * exp(-z * z / 2.);
  }
  return pow(et / .001, -.66666666666666663) * 193.21556 * (pow(et, 1.663753) * .52876731 + 1.) * exp(-et) *
         gaufact * (resfact + 1.);
}

/* diskbb photons in the bins of ear[0..ne] (f_xsdskb__ with f_mcdspc__ at
 * Rin2 = 1, src/donthcomp.c:117-197): five-point Gauss-Legendre per bin. */
static const double GL_W[5] = {.236926885, .47862867, .568888888, .47862867, .236926885};
static const double GL_X[5] = {-.906179846, -.53846931, 0., .53846931, .906179846};

static void diskbb_bins(const double *ear, int ne, double tin, double *photar) {
  for (int i = 1; i <= ne; i++) {
    const double xn = (ear[i] - ear[i - 1]) / 2.0;
    const double xh = xn + ear[i - 1];
    double s = 0.0;
    for (int k = 0; k < 5; k++) {
      const double e = xn * GL_X[k] + xh;
      const double photon = (tin == 0.) ? 0. : orc_mcd_shape(e / tin) * tin * tin * 1. / 361.;
      s += GL_W[k] * photon;
    }
    photar[i - 1] = s * xn;
  }
}

/* One Kompaneets solve (f_thdscompton__ with f_thermlc__,
 * src/donthcomp.c:200-301, :468-648) for the disc seed of temperature tempbb
 * and the electron temperature theta (both in m_e c^2) at photon index gamma.
 * Out: x[0..jmax] the photon energies (m_e c^2), spt[0..jmax-1] the escaping
 * spectrum E F_E (spt[jmax - 1] = 0); returns jmax. */
int orc_nthcomp_solve(double tempbb, double theta, double gamma, double *x, double *spt) {
  static const double delta = .02;
  double c2[NTH_MAX], bet[NTH_MAX], rel[NTH_MAX], dphdot[NTH_MAX], dphesc[NTH_MAX];
  double a[NTH_MAX], b[NTH_MAX], c[NTH_MAX], d[NTH_MAX], g[NTH_MAX], u[NTH_MAX], gam[NTH_MAX], alp[NTH_MAX];
  double ear[NTH_MAX], ph[NTH_MAX];
  memset(c2, 0, sizeof c2);
  memset(bet, 0, sizeof bet);
  memset(rel, 0, sizeof rel);
  memset(dphdot, 0, sizeof dphdot);
  memset(dphesc, 0, sizeof dphesc);
  for (int k = 0; k < NTH_MAX; k++) spt[k] = 0.;
  /* Thomson depth from the photon index (:503) */
  const double gh = gamma + .5;
  const double tautom = sqrt(3. / (theta * (gh * gh - 2.25)) + 2.25) - 1.5;
  const double deltal = delta * log(10.);
  const double xmin = tempbb * 1e-4;
  const double xmax = theta * 40.;
  int jmax = (int)(lg10(xmax / xmin) / delta) + 1;
  if (jmax > 899) jmax = 899;
  /* x_j = xmin 10^(j delta), j = 0..jmax (:534-538) */
  for (int j = 0; j <= jmax; j++) x[j] = xmin * pow(10., j * delta);
  /* Cooper's coefficient at x_{j+1/2} and the Klein-Nishina ratio (:544-566) */
  for (int j = 0; j < jmax; j++) {
    const double w = x[j];
    const double w1 = sqrt(x[j] * x[j + 1]);
    c2[j] = pow(w1, 4.) / (w1 * 4.6 + 1. + w1 * 1.1 * w1);
    if (w <= .05) {
      rel[j] = 1 - w * 2 + w * 26 * w / 5;
    } else {
      const double z1 = (w + 1) / (w * (w * w));
      const double z2 = w * 2 + 1;
      const double z3 = log(z2);
      const double z4 = w * 2 * (w + 1) / z2;
      const double z5 = z3 / 2 / w;
      const double z6 = (w * 3 + 1) / z2 / z2;
      rel[j] = (z1 * (z4 - z3) + z5 - z6) * .75;
    }
  }
  /* the disc seed on the bin centres sqrt(x_j x_{j+1}) up to 50 tempbb
   * (:570-592): dphdot_j = photons / dE, dphdot_0 = dphdot_1 */
  int jmaxth = (int)(lg10(tempbb * 50. / xmin) / delta);
  if (jmaxth > 900) jmaxth = 900;
  if (jmaxth > jmax) jmaxth = jmax;
  for (int j = 0; j < jmaxth - 1; j++) ear[j] = sqrt(x[j] * x[j + 1]) * 511.;
  const int ne = jmaxth - 2;
  diskbb_bins(ear, ne, tempbb * 511., ph);
  for (int j = 1; j <= ne; j++) dphdot[j] = ph[j - 1] * 511. / (ear[j] - ear[j - 1]);
  dphdot[0] = dphdot[1];
  /* escape probability per Thomson time, with the transition between x = 0.1
   * and 1 (:598-626); 1-based bounds jnr, jrel as in the reference */
  int jnr = (int)(lg10(.1 / xmin) / delta + 1);
  if (jnr > jmax - 1) jnr = jmax - 1;
  int jrel = (int)(lg10(1. / xmin) / delta + 1);
  if (jrel > jmax) jrel = jmax;
  const double xnr = x[jnr - 1], xr = x[jrel - 1];
  for (int j = 1; j <= jnr - 1; j++) {
    const double taukn = tautom * rel[j - 1];
    bet[j - 1] = 1 / tautom / (taukn / 3 + 1);
  }
  for (int j = jnr; j <= jrel; j++) {
    const double taukn = tautom * rel[j - 1];
    const double arg = (x[j - 1] - xnr) / (xr - xnr);
    const double flz = 1 - arg;
    bet[j - 1] = 1 / tautom / (taukn / 3 * flz + 1);
  }
  for (int j = jrel + 1; j <= jmax; j++) bet[j - 1] = 1 / tautom;
  /* thermlc (:200-301): a_j u_{j+1} + b_j u_j + c_j u_{j-1} = d_j for the
   * interior points (0-based j = 1 .. jmax-2), zero flux at both ends */
  const double c20 = tautom / deltal;
  for (int j = 1; j <= jmax - 2; j++) {
    const double w1 = sqrt(x[j] * x[j + 1]);
    const double w2 = sqrt(x[j - 1] * x[j]);
    a[j] = -c20 * c2[j] * (theta / deltal / w1 + .5);
    const double t1 = -c20 * c2[j] * (.5 - theta / deltal / w1);
    const double t2 = c20 * c2[j - 1] * (theta / deltal / w2 + .5);
    const double t3 = pow(x[j], 3.) * (tautom * bet[j]);
    b[j] = t1 + t2 + t3;
    c[j] = c20 * c2[j - 1] * (.5 - theta / deltal / w2);
    d[j] = x[j] * dphdot[j];
  }
  const double x32 = sqrt(x[0] * x[1]);
  const double aa = (theta / deltal / x32 + .5) / (theta / deltal / x32 - .5);
  u[jmax - 1] = 0.;
  /* forward elimination with u_0 = aa u_1, back substitution (:258-287) */
  alp[1] = b[1] + c[1] * aa;
  gam[1] = a[1] / alp[1];
  for (int j = 2; j <= jmax - 2; j++) {
    alp[j] = b[j] - c[j] * gam[j - 1];
    gam[j] = a[j] / alp[j];
  }
  g[1] = d[1] / alp[1];
  for (int j = 2; j <= jmax - 3; j++) g[j] = (d[j] - c[j] * g[j - 1]) / alp[j];
  g[jmax - 2] = (d[jmax - 2] - a[jmax - 2] * u[jmax - 1] - c[jmax - 2] * g[jmax - 3]) / alp[jmax - 2];
  u[jmax - 2] = g[jmax - 2];
  for (int jj = jmax - 3; jj >= 1; jj--) u[jj] = g[jj] - gam[jj] * u[jj + 1];
  u[0] = aa * u[1];
  for (int j = 0; j < jmax; j++) dphesc[j] = x[j] * x[j] * u[j] * bet[j] * tautom;
  /* the escaping spectrum E F_E (:630-636) */
  for (int j = 0; j < jmax - 1; j++) spt[j] = dphesc[j] * (x[j] * x[j]);
  return jmax;
}

/* E F_E of the solution at x (f_spp__, src/donthcomp.c:651-681): linear in
 * the bracket [x_{ih-1}, x_ih], ih the first index >= 1 with x_ih >= x
 * (at most nth - 1). */
static double spp_at(double xx, const double *x, int nth, const double *spt) {
  int ih = 1;
  while (ih < nth - 1 && xx > x[ih]) ih++;
  const int il = ih - 1;
  return spt[il] + (spt[ih] - spt[il]) * (xx - x[il]) / (x[ih] - x[il]);
}

/* Photons per bin of ear[0..ne] for the solution (x, spt, nth) seen at
 * redshift z (c_donthcomp's output stage, src/donthcomp.c:735-792): E F_E
 * interpolated at the edges (zero above the grid), divided by E^2,
 * trapezoid per bin, normalised so that E F_E = 1 at 1 keV in the source. */
void orc_nthcomp_photar(const double *x, const double *spt, int nth, const double *ear, int ne, double z,
                        double *photar) {
  const double xn = (z + 1) / 511.;
  const double normfac = 1 / spp_at(1 / (1 / xn), x, nth, spt);
  double prim_prev = 0.;
  int j = 1;  /* 1-based walk of the reference */
  for (int i = 0; i <= ne; i++) {
    while (j <= nth && x[j - 1] * 511. < ear[i] * (z + 1)) j++;
    double prim = 0.;
    if (j <= nth) {
      if (j > 1) {
        const int jl = j - 1;
        prim = spt[jl - 1] + (ear[i] / 511. * (z + 1) - x[jl - 1]) * (spt[jl] - spt[jl - 1]) / (x[jl] - x[jl - 1]);
      } else {
        prim = spt[0];
      }
    }
    if (i > 0)
      photar[i - 1] = (prim / pow(ear[i], 2.) + prim_prev / pow(ear[i - 1], 2.)) * .5 * (ear[i] - ear[i - 1]) * normfac;
    prim_prev = prim;
  }
}

/* spec_nthcomp (src/Xillspec.cpp:243-253): photon index gam, kTe in keV
 * (source frame), seed kT_bb = 0.05 keV disc, energy shift source ->
 * observer ener_shift (z = 1 / ener_shift - 1). */
void orc_spec_nthcomp(double *flux, const double *ener, int n, double kte, double gam, double ener_shift) {
  double x[NTH_MAX + 1], spt[NTH_MAX];
  const int nth = orc_nthcomp_solve(.05 / 511., kte / 511., gam, x, spt);
  orc_nthcomp_photar(x, spt, nth, ener, n, 1 / ener_shift - 1, flux);
}
