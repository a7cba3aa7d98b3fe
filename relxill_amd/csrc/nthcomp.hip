// nthcomp.hip -- the nthcomp primary continuum on the GPU.
//
// spec_nthcomp (src/Xillspec.cpp:243-253): thermal Comptonization of a
// disc-blackbody seed (kT_bb = 0.05 keV, get_nthcomp_param,
// src/relutility.c:625-632) as computed by c_donthcomp (src/donthcomp.c:
// 683-793).  Two kernels:
//   k_nthcomp        one 256-thread workgroup per (Gamma, kTe): the Kompaneets
//                    equation on the log photon grid x_j = 1e-4 T_bb 10^(0.02 j)
//                    (<= 899 points) -- grid, Cooper coefficients, escape
//                    probabilities, the disc seed (five-point Gauss per seed
//                    bin) and the tridiagonal coefficients in parallel, the
//                    tridiagonal solve (forward elimination and back
//                    substitution) by one thread -- written to HBM as
//                    (x, E F_E, jmax);
//   k_nthcomp_bins   one workgroup per spectrum: E F_E interpolated at every
//                    edge of the caller grid (the edge's bracket by binary
//                    search, which finds the bracket of the reference's
//                    monotone walk), divided by E^2, trapezoid per bin, times
//                    the normalisation E F_E(1 keV, source) = 1.
// Every operation keeps the reference's order (contraction off), so results
// differ from the FP64 oracle only by the last-ulp differences of the device
// pow / exp / log.
#include <hip/hip_runtime.h>

#include "engine.h"
#include "rx_math.h"

namespace rx {

namespace {

constexpr int NTH_THREADS = 256;
constexpr int NTH_N = NTH_MAX_POINTS;  // photon energies (the reference's arrays of 900)

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

__device__ inline double lg10(double x) {
#pragma clang fp contract(off)
  return 0.43429448190325182765 * log(x);  // d_lg10, src/donthcomp.c:41-45
}

__device__ inline double mcd_shape(double et) {
#pragma clang fp contract(off)
  const double loget = lg10(et);
  double pos = (loget - lg10(0.001)) / .06 + 1;
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
  q.jmax = min(899, (int)(lg10(xmax / q.xmin) / delta) + 1);
  q.jmaxth = min(900, (int)(lg10(tempbb * 50. / q.xmin) / delta));
  if (q.jmaxth > q.jmax) q.jmaxth = q.jmax;
  q.ne = q.jmaxth - 2;
  q.jnr = min((int)(lg10(.1 / q.xmin) / delta + 1), q.jmax - 1);
  q.jrel = min((int)(lg10(1. / q.xmin) / delta + 1), q.jmax);
  return q;
}

// One Kompaneets solve per workgroup (f_thdscompton__ with f_thermlc__,
// src/donthcomp.c:200-301, :468-648); 0-based arrays, the reference's
// 1-based bounds jnr / jrel kept as they are.
__global__ __launch_bounds__(NTH_THREADS) void k_nthcomp(const NthJob *__restrict__ jobs, double *__restrict__ sol_x,
                                                         double *__restrict__ sol_s, int *__restrict__ sol_n) {
#pragma clang fp contract(off)
  __shared__ double s_x[NTH_N + 1], s_c2[NTH_N], s_bet[NTH_N], s_dph[NTH_N];
  __shared__ double s_a[NTH_N], s_b[NTH_N], s_c[NTH_N], s_d[NTH_N];
  const int tid = threadIdx.x;
  const NthJob jb = jobs[blockIdx.x];
  const double tempbb = jb.ktbb / 511., theta = jb.kte / 511., gamma = jb.gam;
  const NthGrid q = nth_grid(tempbb, theta, gamma);
  const int jmax = q.jmax;
  const double tautom = q.tautom, deltal = q.deltal;
  for (int j = tid; j <= jmax; j += NTH_THREADS) s_x[j] = q.xmin * pow(10., j * .02);
  __syncthreads();
  // Cooper's coefficient at x_{j+1/2}, the escape probability with the
  // Klein-Nishina ratio (:544-566, :598-626)
  const double xnr = s_x[q.jnr - 1], xr = s_x[q.jrel - 1];
  for (int j = tid; j < jmax; j += NTH_THREADS) {
    const double w = s_x[j];
    const double w1 = sqrt(s_x[j] * s_x[j + 1]);
    s_c2[j] = pow(w1, 4.) / (w1 * 4.6 + 1. + w1 * 1.1 * w1);
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
    s_bet[j] = bet;
    s_dph[j] = 0.0;
  }
  // the disc seed (:570-592): bin centres in s_a, photons per bin in s_b
  for (int j = tid; j < q.jmaxth - 1; j += NTH_THREADS) s_a[j] = sqrt(s_x[j] * s_x[j + 1]) * 511.;
  __syncthreads();
  const double tin = tempbb * 511.;
  for (int i = 1 + tid; i <= q.ne; i += NTH_THREADS) {
    const double xn = (s_a[i] - s_a[i - 1]) / 2.0;
    const double xh = xn + s_a[i - 1];
    double ph = 0.0;
    for (int k = 0; k < 5; k++) {
      const double e = xn * c_gl_x[k] + xh;
      const double photon = (tin == 0.) ? 0. : mcd_shape(e / tin) * tin * tin * 1. / 361.;
      ph += c_gl_w[k] * photon;
    }
    s_b[i - 1] = ph * xn;
  }
  __syncthreads();
  for (int j = 1 + tid; j <= q.ne; j += NTH_THREADS) s_dph[j] = s_b[j - 1] * 511. / (s_a[j] - s_a[j - 1]);
  __syncthreads();
  if (tid == 0) s_dph[0] = s_dph[1];
  __syncthreads();
  // tridiagonal coefficients of the interior points (:239-255)
  const double c20 = tautom / deltal;
  for (int j = 1 + tid; j <= jmax - 2; j += NTH_THREADS) {
    const double w1 = sqrt(s_x[j] * s_x[j + 1]);
    const double w2 = sqrt(s_x[j - 1] * s_x[j]);
    s_a[j] = -c20 * s_c2[j] * (theta / deltal / w1 + .5);
    const double t1 = -c20 * s_c2[j] * (.5 - theta / deltal / w1);
    const double t2 = c20 * s_c2[j - 1] * (theta / deltal / w2 + .5);
    const double t3 = pow(s_x[j], 3.) * (tautom * s_bet[j]);
    s_b[j] = t1 + t2 + t3;
    s_c[j] = c20 * s_c2[j - 1] * (.5 - theta / deltal / w2);
    s_d[j] = s_x[j] * s_dph[j];
  }
  __syncthreads();
  // the solve (:258-287), in place: s_b <- alp, s_a <- gam, s_d <- g, then u
  if (tid == 0) {
    const double x32 = sqrt(s_x[0] * s_x[1]);
    const double aa = (theta / deltal / x32 + .5) / (theta / deltal / x32 - .5);
    s_b[1] = s_b[1] + s_c[1] * aa;
    s_a[1] = s_a[1] / s_b[1];
    for (int j = 2; j <= jmax - 2; j++) {
      s_b[j] = s_b[j] - s_c[j] * s_a[j - 1];
      s_a[j] = s_a[j] / s_b[j];
    }
    s_d[1] = s_d[1] / s_b[1];
    for (int j = 2; j <= jmax - 3; j++) s_d[j] = (s_d[j] - s_c[j] * s_d[j - 1]) / s_b[j];
    // the last row reads a_{jmax-2} before it became gam: recompute it
    const double w1 = sqrt(s_x[jmax - 2] * s_x[jmax - 1]);
    const double a_last = -c20 * s_c2[jmax - 2] * (theta / deltal / w1 + .5);
    const double u_top = 0.;
    s_d[jmax - 2] = (s_d[jmax - 2] - a_last * u_top - s_c[jmax - 2] * s_d[jmax - 3]) / s_b[jmax - 2];
    s_d[jmax - 1] = u_top;
    for (int jj = jmax - 3; jj >= 1; jj--) s_d[jj] = s_d[jj] - s_a[jj] * s_d[jj + 1];
    s_d[0] = aa * s_d[1];
  }
  __syncthreads();
  // escaping photons and E F_E (:297-300, :630-636)
  double *xo = sol_x + (size_t)blockIdx.x * (NTH_N + 1);
  double *so = sol_s + (size_t)blockIdx.x * NTH_N;
  for (int j = tid; j <= jmax; j += NTH_THREADS) xo[j] = s_x[j];
  for (int j = tid; j < NTH_N; j += NTH_THREADS) {
    double v = 0.0;
    if (j < jmax - 1) {
      const double dphesc = s_x[j] * s_x[j] * s_d[j] * s_bet[j] * tautom;
      v = dphesc * (s_x[j] * s_x[j]);
    }
    so[j] = v;
  }
  if (tid == 0) sol_n[blockIdx.x] = jmax;
}

}  // namespace

// E F_E of a solution at x (f_spp__, src/donthcomp.c:651-681): the first
// index ih >= 1 with x_ih >= x (at most nth - 1), linear in [x_{ih-1}, x_ih].
__device__ double nthcomp_spp(const double *x, const double *spt, int nth, double xx) {
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
__device__ double nthcomp_edge(const double *x, const double *spt, int nth, double ear, double z) {
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

// photons in the bin [e0, e1] from the edge values p0, p1 (E F_E) with the
// normalisation normfac
__device__ double nthcomp_bin(double e0, double e1, double p0, double p1, double normfac) {
#pragma clang fp contract(off)
  return (p1 / (e1 * e1) + p0 / (e0 * e0)) * .5 * (e1 - e0) * normfac;
}

namespace {

__global__ __launch_bounds__(NTH_THREADS) void k_nthcomp_bins(const double *__restrict__ ener, int n,
                                                              const double *__restrict__ shift,
                                                              const double *__restrict__ sol_x,
                                                              const double *__restrict__ sol_s,
                                                              const int *__restrict__ sol_n, double *__restrict__ out) {
#pragma clang fp contract(off)
  const int s = blockIdx.x;
  const double *x = sol_x + (size_t)s * (NTH_N + 1);
  const double *spt = sol_s + (size_t)s * NTH_N;
  const int nth = sol_n[s];
  const double z = 1 / shift[s] - 1;
  const double xn = (z + 1) / 511.;
  const double normfac = 1 / nthcomp_spp(x, spt, nth, 1 / (1 / xn));
  for (int i = threadIdx.x; i < n; i += NTH_THREADS) {
    const double p0 = nthcomp_edge(x, spt, nth, ener[i], z);
    const double p1 = nthcomp_edge(x, spt, nth, ener[i + 1], z);
    out[(size_t)s * n + i] = nthcomp_bin(ener[i], ener[i + 1], p0, p1, normfac);
  }
}

}  // namespace

int launch_nthcomp(const NthJob *jobs, int n_jobs, double *sol_x, double *sol_s, int *sol_n, void *stream) {
  if (n_jobs <= 0) return 0;
  hipLaunchKernelGGL(k_nthcomp, dim3(n_jobs), dim3(NTH_THREADS), 0, (hipStream_t)stream, jobs, sol_x, sol_s,
                     sol_n);
  return (int)hipGetLastError();
}

// rx_nthcomp: n_spec spectra on one caller grid (host arrays in and out).
int nthcomp_spectra(const double *ener, int n_flux, int n_spec, const double *kte, const double *gam,
                    const double *ener_shift, double *flux, std::string &err) {
  for (int s = 0; s < n_spec; s++) {
    // the solve's grid needs kTe well above the seed and a finite Thomson depth
    if (!(kte[s] >= 0.1 && kte[s] <= 1e4) || !(gam[s] > 1.0 && gam[s] < 10.0) ||
        !(ener_shift[s] > 0.0 && ener_shift[s] < 1e3)) {
      err = "nthcomp parameters out of range (kTe 0.1-1e4 keV, 1 < Gamma < 10, energy shift > 0)";
      return RX_ERR_ARG;
    }
  }
  for (int k = 0; k < n_flux; k++)
    if (!(ener[k + 1] > ener[k]) || !(ener[k] > 0.0)) {
      err = "energy grid must be positive and ascending";
      return RX_ERR_ARG;
    }
  std::vector<NthJob> jobs(n_spec);
  for (int s = 0; s < n_spec; s++) jobs[s] = NthJob{kte[s], gam[s], 0.05};
  DevBuf d_jobs, d_e, d_shift, d_x, d_s, d_n, d_out;
  if (d_jobs.ensure(sizeof(NthJob) * n_spec) || d_e.ensure(sizeof(double) * (n_flux + 1)) ||
      d_shift.ensure(sizeof(double) * n_spec) || d_x.ensure(sizeof(double) * (NTH_N + 1) * n_spec) ||
      d_s.ensure(sizeof(double) * NTH_N * n_spec) || d_n.ensure(sizeof(int) * n_spec) ||
      d_out.ensure(sizeof(double) * (size_t)n_flux * n_spec)) {
    err = "device allocation failed";
    return RX_ERR_NOMEM;
  }
  int e = (int)hipMemcpy(d_jobs.p, jobs.data(), sizeof(NthJob) * n_spec, hipMemcpyHostToDevice);
  e |= (int)hipMemcpy(d_e.p, ener, sizeof(double) * (n_flux + 1), hipMemcpyHostToDevice);
  e |= (int)hipMemcpy(d_shift.p, ener_shift, sizeof(double) * n_spec, hipMemcpyHostToDevice);
  e |= launch_nthcomp(d_jobs.as<NthJob>(), n_spec, d_x.as<double>(), d_s.as<double>(), d_n.as<int>(), nullptr);
  hipLaunchKernelGGL(k_nthcomp_bins, dim3(n_spec), dim3(NTH_THREADS), 0, nullptr, d_e.as<double>(), n_flux,
                     d_shift.as<double>(), d_x.as<double>(), d_s.as<double>(), d_n.as<int>(), d_out.as<double>());
  e |= (int)hipGetLastError();
  e |= (int)hipMemcpy(flux, d_out.p, sizeof(double) * (size_t)n_flux * n_spec, hipMemcpyDeviceToHost);
  if (e) {
    (void)hipGetLastError();
    err = "device error in the nthcomp kernels";
    return RX_ERR_DEVICE;
  }
  return RX_OK;
}

}  // namespace rx
