// rx_nthcomp.h -- device building blocks of the nthcomp primary continuum
// (c_donthcomp, src/donthcomp.c:683-793; spec_nthcomp, src/Xillspec.cpp:243-253),
// shared by the rx_nthcomp kernels and the relxill primary of the nthcomp
// models.  Every operation keeps the reference's order (contraction off).
#pragma once

#include <hip/hip_runtime.h>

#include "rx_engine.h"

namespace rx {

constexpr int NTH_N = NTH_MAX_POINTS;  // photon energies (the reference's arrays of 900)

namespace {

// diskbb interpolation formula (f_mcdint__, src/donthcomp.c:50-115): the
// model's published coefficients
__constant__ double c_mcd_gc[3] = {.078196667, -1.066202, 1.192418};
__constant__ double c_mcd_gw[3] = {.5207874, .513457, .4077983};
__constant__ double c_mcd_gn[3] = {.3728691, .039775528, .037766505};
__constant__ double c_mcd_res[98] = {
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
// five-point Gauss-Legendre weights and nodes of f_xsdskb__ (:149-151)
__constant__ double c_gl_w[5] = {.236926885, .47862867, .568888888, .47862867, .236926885};
__constant__ double c_gl_x[5] = {-.906179846, -.53846931, 0., .53846931, .906179846};

__device__ inline double nth_lg10(double x) {
#pragma clang fp contract(off)
  return 0.43429448190325182765 * log(x);  // d_lg10, src/donthcomp.c:41-45
}

__device__ inline double nth_mcd_shape(double et) {
#pragma clang fp contract(off)
  const double loget = nth_lg10(et);
  double pos = (loget - nth_lg10(0.001)) / .06 + 1;
  const int j = (int)pos;
  double resfact;
  if (j < 1) {
    resfact = c_mcd_res[0];
  } else if (j >= 98) {
    resfact = c_mcd_res[97];
  } else {
    pos -= j;
    resfact = c_mcd_res[j - 1] * (1. - pos) + c_mcd_res[j] * pos;
  }
  double gaufact = 1.;
  for (int k = 0; k < 3; k++) {
    const double z = (loget - c_mcd_gc[k]) / c_mcd_gw[k];
    gaufact += c_mcd_gn[k] * exp(-z * z / 2.);
  }
  return pow(et / .001, -.66666666666666663) * 193.21556 * (pow(et, 1.663753) * .52876731 + 1.) * exp(-et) *
         gaufact * (resfact + 1.);
}

// the scalars every thread of a solve derives the same way
struct NthGrid {
  double tautom, deltal, xmin;
  int jmax, jmaxth, ne, jnr, jrel;
};

__device__ inline NthGrid nth_grid(double tempbb, double theta, double gamma) {
#pragma clang fp contract(off)
  NthGrid q;
  const double delta = .02;
  const double gh = gamma + .5;
  q.tautom = sqrt(3. / (theta * (gh * gh - 2.25)) + 2.25) - 1.5;
  q.deltal = delta * log(10.);
  q.xmin = tempbb * 1e-4;
  const double xmax = theta * 40.;
  q.jmax = min(899, (int)(nth_lg10(xmax / q.xmin) / delta) + 1);
  q.jmaxth = min(900, (int)(nth_lg10(tempbb * 50. / q.xmin) / delta));
  if (q.jmaxth > q.jmax) q.jmaxth = q.jmax;
  q.ne = q.jmaxth - 2;
  q.jnr = min((int)(nth_lg10(.1 / q.xmin) / delta + 1), q.jmax - 1);
  q.jrel = min((int)(nth_lg10(1. / q.xmin) / delta + 1), q.jmax);
  return q;
}

}  // namespace

// The LDS of one Kompaneets solve (57.6 KB).
struct NthLds {
  double x[NTH_N + 1], c2[NTH_N], bet[NTH_N], dph[NTH_N];
  double a[NTH_N], b[NTH_N], c[NTH_N], d[NTH_N];
};

// One Kompaneets solve by the nt threads of a workgroup (f_thdscompton__ with
// f_thermlc__, src/donthcomp.c:200-301, :468-648; 0-based arrays, the
// reference's 1-based bounds jnr / jrel kept as they are): the log photon
// grid x_j = 1e-4 T_bb 10^(0.02 j), Cooper's coefficients, the escape
// probabilities, the disc seed (five-point Gauss per seed bin) and the
// tridiagonal coefficients in parallel, the solve (forward elimination and
// back substitution) by thread 0.  Leaves x[0..jmax] in L.x and E F_E in
// L.c[0..NTH_N) (0 from jmax - 1 on); returns jmax.  Begins and ends with a
// barrier, so L may hold other data before the call.
__device__ inline int nth_solve(NthLds &L, double kte, double gam, double ktbb, int tid, int nt) {
#pragma clang fp contract(off)
  const double tempbb = ktbb / 511., theta = kte / 511., gamma = gam;
  const NthGrid q = nth_grid(tempbb, theta, gamma);
  const int jmax = q.jmax;
  const double tautom = q.tautom, deltal = q.deltal;
  __syncthreads();
  for (int j = tid; j <= jmax; j += nt) L.x[j] = q.xmin * pow(10., j * .02);
  __syncthreads();
  // Cooper's coefficient at x_{j+1/2}, the escape probability with the
  // Klein-Nishina ratio (:544-566, :598-626)
  const double xnr = L.x[q.jnr - 1], xr = L.x[q.jrel - 1];
  for (int j = tid; j < jmax; j += nt) {
    const double w = L.x[j];
    const double w1 = sqrt(L.x[j] * L.x[j + 1]);
    L.c2[j] = pow(w1, 4.) / (w1 * 4.6 + 1. + w1 * 1.1 * w1);
    double rel;
    if (w <= .05) {
      rel = 1 - w * 2 + w * 26 * w / 5;
    } else {
      const double z1 = (w + 1) / (w * (w * w));
      const double z2 = w * 2 + 1;
      const double z3 = log(z2);
      const double z4 = w * 2 * (w + 1) / z2;
      const double z5 = z3 / 2 / w;
      const double z6 = (w * 3 + 1) / z2 / z2;
      rel = (z1 * (z4 - z3) + z5 - z6) * .75;
    }
    const int j1 = j + 1;  // the reference's 1-based index
    const double taukn = tautom * rel;
    double bet;
    if (j1 <= q.jnr - 1) {
      bet = 1 / tautom / (taukn / 3 + 1);
    } else if (j1 <= q.jrel) {
      const double arg = (w - xnr) / (xr - xnr);
      const double flz = 1 - arg;
      bet = 1 / tautom / (taukn / 3 * flz + 1);
    } else {
      bet = 1 / tautom;
    }
    L.bet[j] = bet;
    L.dph[j] = 0.0;
  }
  // the disc seed (:570-592): bin centres in a, photons per bin in b
  for (int j = tid; j < q.jmaxth - 1; j += nt) L.a[j] = sqrt(L.x[j] * L.x[j + 1]) * 511.;
  __syncthreads();
  const double tin = tempbb * 511.;
  for (int i = 1 + tid; i <= q.ne; i += nt) {
    const double xn = (L.a[i] - L.a[i - 1]) / 2.0;
    const double xh = xn + L.a[i - 1];
    double ph = 0.0;
    for (int k = 0; k < 5; k++) {
      const double e = xn * c_gl_x[k] + xh;
      const double photon = (tin == 0.) ? 0. : nth_mcd_shape(e / tin) * tin * tin * 1. / 361.;
      ph += c_gl_w[k] * photon;
    }
    L.b[i - 1] = ph * xn;
  }
  __syncthreads();
  for (int j = 1 + tid; j <= q.ne; j += nt) L.dph[j] = L.b[j - 1] * 511. / (L.a[j] - L.a[j - 1]);
  __syncthreads();
  if (tid == 0) L.dph[0] = L.dph[1];
  __syncthreads();
  // tridiagonal coefficients of the interior points (:239-255)
  const double c20 = tautom / deltal;
  for (int j = 1 + tid; j <= jmax - 2; j += nt) {
    const double w1 = sqrt(L.x[j] * L.x[j + 1]);
    const double w2 = sqrt(L.x[j - 1] * L.x[j]);
    L.a[j] = -c20 * L.c2[j] * (theta / deltal / w1 + .5);
    const double t1 = -c20 * L.c2[j] * (.5 - theta / deltal / w1);
    const double t2 = c20 * L.c2[j - 1] * (theta / deltal / w2 + .5);
    const double t3 = pow(L.x[j], 3.) * (tautom * L.bet[j]);
    L.b[j] = t1 + t2 + t3;
    L.c[j] = c20 * L.c2[j - 1] * (.5 - theta / deltal / w2);
    L.d[j] = L.x[j] * L.dph[j];
  }
  __syncthreads();
  // the solve (:258-287), in place: b <- alp, a <- gam, d <- g, then u
  if (tid == 0) {
    const double x32 = sqrt(L.x[0] * L.x[1]);
    const double aa = (theta / deltal / x32 + .5) / (theta / deltal / x32 - .5);
    L.b[1] = L.b[1] + L.c[1] * aa;
    L.a[1] = L.a[1] / L.b[1];
    for (int j = 2; j <= jmax - 2; j++) {
      L.b[j] = L.b[j] - L.c[j] * L.a[j - 1];
      L.a[j] = L.a[j] / L.b[j];
    }
    L.d[1] = L.d[1] / L.b[1];
    for (int j = 2; j <= jmax - 3; j++) L.d[j] = (L.d[j] - L.c[j] * L.d[j - 1]) / L.b[j];
    // the last row reads a_{jmax-2} before it became gam: recompute it
    const double w1 = sqrt(L.x[jmax - 2] * L.x[jmax - 1]);
    const double a_last = -c20 * L.c2[jmax - 2] * (theta / deltal / w1 + .5);
    const double u_top = 0.;
    L.d[jmax - 2] = (L.d[jmax - 2] - a_last * u_top - L.c[jmax - 2] * L.d[jmax - 3]) / L.b[jmax - 2];
    L.d[jmax - 1] = u_top;
    for (int jj = jmax - 3; jj >= 1; jj--) L.d[jj] = L.d[jj] - L.a[jj] * L.d[jj + 1];
    L.d[0] = aa * L.d[1];
  }
  __syncthreads();
  // escaping photons and E F_E (:297-300, :630-636), into c
  for (int j = tid; j < NTH_N; j += nt) {
    double v = 0.0;
    if (j < jmax - 1) {
      const double dphesc = L.x[j] * L.x[j] * L.d[j] * L.bet[j] * tautom;
      v = dphesc * (L.x[j] * L.x[j]);
    }
    L.c[j] = v;
  }
  __syncthreads();
  return jmax;
}

// E F_E of a solution at x (f_spp__, src/donthcomp.c:651-681): the first
// index ih >= 1 with x_ih >= x (at most nth - 1), linear in [x_{ih-1}, x_ih].
__device__ inline double nthcomp_spp(const double *x, const double *spt, int nth, double xx) {
#pragma clang fp contract(off)
  int lo = 1, hi = nth - 1;  // first ih in [1, nth-1] with !(xx > x[ih]), else nth - 1
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (xx > x[mid]) lo = mid + 1;
    else hi = mid;
  }
  const int ih = lo, il = ih - 1;
  return spt[il] + (spt[ih] - spt[il]) * (xx - x[il]) / (x[ih] - x[il]);
}

// E F_E at the grid edge ear (keV) for redshift z (c_donthcomp, :757-774):
// the reference walks j up while x_{j-1} 511 < ear (z + 1); the bracket is
// the first such failure, found here by binary search.
__device__ inline double nthcomp_edge(const double *x, const double *spt, int nth, double ear, double z) {
#pragma clang fp contract(off)
  int lo = 1, hi = nth + 1;  // first j in [1, nth] with !(x[j-1] 511 < ear (z+1)), else nth + 1
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (x[mid - 1] * 511. < ear * (z + 1)) lo = mid + 1;
    else hi = mid;
  }
  const int j = lo;
  if (j > nth) return 0.0;
  if (j == 1) return spt[0];
  const int jl = j - 1;
  return spt[jl - 1] + (ear / 511. * (z + 1) - x[jl - 1]) * (spt[jl] - spt[jl - 1]) / (x[jl] - x[jl - 1]);
}

// the normalisation of c_donthcomp (:747-756): E F_E(1 keV, source) = 1
__device__ inline double nthcomp_normfac(const double *x, const double *spt, int nth, double z) {
#pragma clang fp contract(off)
  const double xn = (z + 1) / 511.;
  return 1 / nthcomp_spp(x, spt, nth, 1 / (1 / xn));
}

// photons in the bin [e0, e1] from the edge values p0, p1 (E F_E) with the
// normalisation normfac
__device__ inline double nthcomp_bin(double e0, double e1, double p0, double p1, double normfac) {
#pragma clang fp contract(off)
  return (p1 / (e1 * e1) + p0 / (e0 * e0)) * .5 * (e1 - e0) * normfac;
}

}  // namespace rx
