// kernels.hip -- CDNA4 (gfx950) kernels of the relativistic smearing path.
//
//   k_reset      clears the run's counters
//   k_setup      one workgroup per request: (a, mu0) bilinear table
//                interpolation (interpol_relTable, src/Relprofile.cpp:141-303),
//                fine radial grid, emissivity (src/Rellp.cpp:518-564), energy
//                windows, angle distribution (src/Relprofile.cpp:908-938), and
//                the request's work items (radius chunk x bin segment)
//   k_integrate  persistent 64-lane wavefronts over the work items: lane =
//                fine radius, loop over the segment's bins, fixed-order wave
//                sum per bin (calc_relline_profile, src/Relprofile.cpp:857-900)
//   k_fixup      FP64 re-check of FP32 Romberg stop tests in the guard band
//   k_finalize   one workgroup per request: photons per bin and
//                renormalisation (src/Relprofile.cpp:749-796)
//   k_conv       one workgroup per convolution request: rebin onto the
//                convolution grid, real FFT convolution held in LDS with the
//                1 keV rotation and E/dE weighting, band normalisation, rebin
//                back and zeroing outside 0.01-1000 keV
//                (src/Relbase.cpp:93-289)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "rx_engine.h"
#include "rx_math.h"
#include "rx_nthcomp.h"

namespace rx {

namespace {

constexpr int SETUP_THREADS = 256;
constexpr int CONV_THREADS = 256;      // smallest convolution workgroup
constexpr int XILL_THREADS = 256;
constexpr int XILL_BINS = 4;    // table bins per k_xill thread, one-zone requests
constexpr int XILL_BINS_Z = 2;  // the same for requests of LP_ZONES zones (ten accumulators per bin)
constexpr int PRIM_THREADS = 256;
constexpr int CONV_THREADS_MAX = 1024;  // n/8 threads: 512 for 4096 bins, 1024 for 8192

__device__ inline double lp_ifac(float value, const float *arr, int n, int &ind) {
  // get_ipol_factor: float search and float quotient (src/relutility.c:419-423)
  ind = bsearch_asc(arr, n, value);
  return (double)((value - arr[ind]) / (arr[ind + 1] - arr[ind]));
}

// spec_cutoffpl (src/Xillspec.cpp:215-241): photons of the cut-off power law
// exp(1/Ecut) E^-Gamma exp(-E/Ecut) dE in [e0, e1] at E = (e0 + e1) / 2, as
// one exponential of ln_en = ln E (precomputed per grid bin): it differs from
// the product of the three factors by a few ulp of the exponent
__device__ inline double cutoffpl_bin(double e0, double e1, double ln_en, double ecut, double gam) {
  const double en = 0.5 * (e0 + e1);
  return exp(1.0 / ecut - gam * ln_en - en / ecut) * (e1 - e0);
}

// spec_blackbody (src/Xillspec.cpp:255-268): photons of E^2 / (kT^4 (exp(E/kT) - 1))
// dE in [e0, e1] at E = (e0 + e1) / 2
__device__ inline double bbody_bin(double e0, double e1, double kt) {
  const double en = 0.5 * (e0 + e1);
  return en * en / (pow(kt, 4) * (exp(en / kt) - 1)) * (e1 - e0);
}

// 1e15 / (4 pi): the xillver tables' incident flux normalisation (src/Xillspec.cpp:203)
constexpr double NORM_XILLVER_TABLE = 1e15 / 4.0 / 3.14159265358979323846;

// A xillver table value as the reference holds it after renorm_xill_spec
// (src/xilltable.c:478-486): the resident rows are divided by 10^logXi; rows
// of a table without a density axis are also divided by dens_div =
// 10^(logN - 15) (rounded to float, as the reference's float array), unless
// logN is 15 (dens_div 0).
__device__ __forceinline__ double xval(float v, double dens_div) {
  return dens_div != 0.0 ? (double)(float)((double)v / dens_div) : (double)v;
}

// This thread's part of calcNormWrtXillverTableSpec's sum (src/Xillspec.cpp:179-205)
// for the cut-off power law on the 500-bin 0.1-1000 keV grid ce.
__device__ inline double xill_norm_part(const DevConvGrid &cg, double ecut, double gam, int tid, int nt) {
  double sp = 0.0;
  for (int k = tid; k < N_ENER_COARSE; k += nt) {
    const double e0 = cg.coarse[k], e1 = cg.coarse[k + 1];
    if (e0 >= 0.1 && e0 <= 1000.0)
      sp += cutoffpl_bin(e0, e1, cg.coarse_ln[k], ecut, gam) * 0.5 * (e0 + e1) * 1e20 * 1.602177e-09;
  }
  return sp;
}

// The same sum for the blackbody of kT_bb kt.
__device__ inline double bbody_norm_part(const DevConvGrid &cg, double kt, int tid, int nt) {
  double sp = 0.0;
  for (int k = tid; k < N_ENER_COARSE; k += nt) {
    const double e0 = cg.coarse[k], e1 = cg.coarse[k + 1];
    if (e0 >= 0.1 && e0 <= 1000.0) sp += bbody_bin(e0, e1, kt) * 0.5 * (e0 + e1) * 1e20 * 1.602177e-09;
  }
  return sp;
}

struct LpFine {
  double emis, del_emit, del_inc;
  int err;
};

// rebin_emisprofile_on_radial_grid for one fine radius (src/Rellp.cpp:120-175):
// the table walk reduces to kk = clamp(#{jet_rad <= r} - 1, 0, n-2).
__device__ inline LpFine lp_on_fine(double re, const double *jr, const double *je,
                                    const double *jd, const double *ji, int ntr) {
  LpFine o;
  o.err = 0;
  int lo = 0, hi = ntr;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (jr[mid] <= re) lo = mid + 1; else hi = mid;
  }
  int kk = lo - 1;
  if (kk < 0) kk = 0;
  if (kk >= ntr - 1) {
    if (!(re - 1000.0 <= 1e-6)) o.err = 22;
    kk = ntr - 2;
  }
  double inter_r;
  if (jd[kk] / PI * 180.0 <= 75.0) {  // get_ipol_factor_radius, src/relutility.c:407-417
    inter_r = (re - jr[kk]) / (jr[kk + 1] - jr[kk]);
  } else {
    inter_r = (log(re) - log(jr[kk])) / (log(jr[kk + 1]) - log(jr[kk]));
  }
  o.emis = exp(inter_r * log(je[kk + 1]) + (1.0 - inter_r) * log(je[kk]));
  o.del_emit = lerp1(inter_r, jd[kk], jd[kk + 1]);
  o.del_inc = lerp1(inter_r, ji[kk], ji[kk + 1]);
  return o;
}

// g = E/E_i for the lamp post (src/Relphysics.cpp:163-207)
__device__ inline double gi_potential_lp(double r, double a, double h, double bet, double del) {
  double ut_d = ((r * sqrt(r) + a) / (sqrt(r) * sqrt(r * r - 3 * r + 2 * a * sqrt(r))));
  double ut_h = sqrt((h * h + a * a) / (h * h - 2 * h + a * a));
  double gi = ut_d / ut_h;
  if (fabs(bet) < 1e-6) return gi;
  double gam = 1.0 / sqrt(1.0 - bet * bet);
  double sgn = (del > PI / 2) ? -1.0 : 1.0;
  double delta_eq = h * h - 2 * h + a * a;
  double sd = sin(del);
  double hh = h * h + a * a;
  double q2 = (sd * sd) * (hh * hh / delta_eq) - a * a;
  double beta_fac = sqrt(hh * hh - delta_eq * (q2 + a * a));
  beta_fac = gam * (1.0 + sgn * beta_fac / hh * bet);
  return gi / beta_fac;
}

__device__ inline double relat_abberation(double del, double beta) {
  return acos((cos(del) - beta) / (1 - beta * cos(del)));
}

// corrected_gshift_fluxboost_factor (src/Relreturn_Corona.cpp:39-83): the flux
// boost of the xillver reflection under the energy shift g, from its boost
// ratio c to a power law at g = 1.5 (a parabola through 0 for c < 1, a line
// otherwise; at most 1 below g = 1, never negative)
__device__ inline double gshift_boost(double c, double g, double gamma) {
  const double g0 = 2. / 3;
  double corr;
  if (c < 1) {
    const double a = (c / g0 - 1) / (g0 - 1);
    const double b = 1 - a;
    corr = (g >= 1) ? 1. / g * (1. / g * a + b) : g * (g * a + b);
  } else {
    const double alin = (c - 1) / (g0 - 1);
    const double blin = 1 - alin;
    corr = (g >= 1) ? (1. / g * alin + blin) : (g * alin + blin);
  }
  double boost = pow(g, gamma) * corr;
  if (g < 1 && boost > 1) boost = 1;
  if (boost < 0) boost = 0;
  return boost;
}

// interpol_lptable (src/Rellp.cpp:219-247) of one request into the block's
// LDS rows: radius (lerped in spin), intensity, emission and incidence angle
// ((1 - f_a) lerp_h(spin lo) + f_a lerp_h(spin hi), float factors) at the
// table's n_r radii.  Called by every thread of the block (two barriers).
__device__ inline void lp_rows(const DevLpTable &L, const ReqParams &rq, double *jr, double *je, double *jd,
                               double *ji, int tid, int nt) {
  __shared__ int s_i[3];
  __shared__ double s_f[3];
  if (tid == 0) {
    int lia;
    const double lfa = lp_ifac((float)rq.a, L.a, L.n_a, lia);
    int h0, h1;
    const double fh0 = lp_ifac((float)rq.height, L.h + (size_t)lia * L.n_h, L.n_h, h0);
    const double fh1 = lp_ifac((float)rq.height, L.h + (size_t)(lia + 1) * L.n_h, L.n_h, h1);
    s_i[0] = lia;
    s_i[1] = h0;
    s_i[2] = h1;
    s_f[0] = lfa;
    s_f[1] = fh0;
    s_f[2] = fh1;
  }
  __syncthreads();
  const int ntr = L.n_r;
  const int lia = s_i[0], h0 = s_i[1], h1 = s_i[2];
  const double lfa = s_f[0], fh0 = s_f[1], fh1 = s_f[2];
  for (int ii = tid; ii < ntr; ii += nt) {
    jr[ii] = lerp1(lfa, (double)L.rad[(size_t)lia * ntr + ii], (double)L.rad[(size_t)(lia + 1) * ntr + ii]);
    const size_t l0 = ((size_t)lia * L.n_h + h0) * ntr + ii;
    const size_t u0 = ((size_t)lia * L.n_h + h0 + 1) * ntr + ii;
    const size_t l1 = ((size_t)(lia + 1) * L.n_h + h1) * ntr + ii;
    const size_t u1 = ((size_t)(lia + 1) * L.n_h + h1 + 1) * ntr + ii;
    je[ii] = (1.0 - lfa) * lerp1(fh0, L.intens[l0], L.intens[u0]) + (lfa)*lerp1(fh1, L.intens[l1], L.intens[u1]);
    jd[ii] = (1.0 - lfa) * lerp1(fh0, L.del[l0], L.del[u0]) + (lfa)*lerp1(fh1, L.del[l1], L.del[u1]);
    ji[ii] = (1.0 - lfa) * lerp1(fh0, L.del_inc[l0], L.del_inc[u0]) + (lfa)*lerp1(fh1, L.del_inc[l1], L.del_inc[u1]);
  }
  __syncthreads();
}

// One radial zone of a relxill request (IonGradient, src/IonGradient.cpp:
// 182-209, :421-430): its centre, for the lamp post the emission angle there
// (inv_rebin_mean of the fine grid's angles, src/relutility.c:636-663), the
// source -> disk energy shift and the Ecut the zone sees with its xillver
// bracket (float factor, clamps, ensure_ecut_within_boundarys,
// src/xilltable.c:1054-1125).  *ect receives that Ecut (lamp post).
__device__ inline ZoneRec zone_record(const DevXillTable &xt, const ReqParams &rq, int z, bool lp, const double *jr,
                                      const double *je, const double *jd, const double *ji, int ntr, int nr,
                                      double *ect) {
  const double *zg = rq.zgrid;
  ZoneRec o;
  o.rmean = 0.5 * (zg[z] + zg[z + 1]);
  o.del_emit = 0.0;
  o.shift = 1.0;
  o.ncf = 1.0;
  o.ect_ind = xt.ect_dim >= 0 ? rq.xind[xt.ect_dim] : 0;
  o.ect_fac = xt.ect_dim >= 0 ? rq.xfac[xt.ect_dim] : 0.0;
  o.lxi_ind = xt.lxi_dim >= 0 ? rq.xind[xt.lxi_dim] : 0;
  o.lxi_fac = xt.lxi_dim >= 0 ? rq.xfac[xt.lxi_dim] : 0.0;
  *ect = rq.xect_src;
  if (rq.ion_grad == 1 && xt.lxi_dim >= 0) {
    // the power-law gradient (calc_ion_grad_pl, src/IonGradient.cpp:174-182)
    // from the input logXi in the zones' mean radius, kept within 0-4.7
    // (adjust_parameter_to_be_within_xillver_bounds, :36-44, :388), and its
    // bracket on the table's logXi axis as the request's parameters get
    // theirs (float value, clamped into the table, float factor)
    const double rm0 = 0.5 * (zg[0] + zg[1]);
    double lxi = exp(rq.xlxi0) * pow(o.rmean / rm0, -1.0 * rq.ion_index);
    lxi = log(lxi);
    if (lxi < 0.0) lxi = 0.0;
    else if (lxi > 4.7) lxi = 4.7;
    const float *vals = xt.lxi_vals;
    const int n = xt.n_lxi;
    float v = (float)lxi;
    int ind = bsearch_asc(vals, n, v);
    if (ind < 0) ind = 0;
    else if (ind > n - 2) ind = n - 2;
    if (v < vals[0]) v = vals[0];
    else if (v > vals[n - 1]) v = vals[n - 1];
    const float num = v - vals[ind];
    const float den = vals[ind + 1] - vals[ind];
    o.lxi_ind = ind;
    o.lxi_fac = (float)(num / den);
  }
  if (lp) {
    int lo = 0, hi = nr - 1;  // re(lo) > rmean >= re(hi)
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (fine_radius(rq.rin, rq.rout, mid, nr) > o.rmean) lo = mid; else hi = mid;
    }
    const double r0 = fine_radius(rq.rin, rq.rout, lo, nr), r1 = fine_radius(rq.rin, rq.rout, lo + 1, nr);
    const LpFine f0 = lp_on_fine(r0, jr, je, jd, ji, ntr);
    const LpFine f1 = lp_on_fine(r1, jr, je, jd, ji, ntr);
    o.del_emit = lerp1((o.rmean - r1) / (r0 - r1), f1.del_emit, f0.del_emit);
    o.shift = gi_potential_lp(o.rmean, rq.a, rq.height, rq.beta, o.del_emit);
    *ect = rq.xect_src * o.shift;
    if (xt.ect_dim >= 0) {
      const float *vals = xt.ect_vals;
      const int n = xt.n_ect;
      float v = (float)*ect;
      int ind = bsearch_asc(vals, n, v);
      if (ind < 0) ind = 0;
      else if (ind > n - 2) ind = n - 2;
      if (v < vals[0]) v = vals[0];
      else if (v > vals[n - 1]) v = vals[n - 1];
      const float num = v - vals[ind];
      const float den = vals[ind + 1] - vals[ind];
      double fac = (float)(num / den);
      if (*ect <= xt.q_vals[0]) fac = 0.0;
      if (*ect >= xt.q_vals[xt.n_q - 1]) fac = 1.0;
      o.ect_ind = ind;
      o.ect_fac = fac;
    }
  }
  return o;
}

// Where a request's profile accumulates: the caller output for line models,
// the convolution-grid profile for convolution models.
__device__ inline double *profile_acc(const BatchDev &b, const ReqParams &rq) {
  return (rq.prof_off >= 0) ? b.prof + rq.prof_off : b.out + rq.flux_off;
}

// Index guess for an ascending grid that is close to log-spaced (every grid
// here is): e0[k] ~ e0[0] exp(k dlog).
struct GridGuess {
  float lg0, scale;  // log2(e0[0]), n0 / log2(e0[n0] / e0[0])
  double first, last;
};

__device__ __forceinline__ GridGuess grid_guess(const double *e0, int n0) {
  GridGuess g;
  g.first = e0[0];
  g.last = e0[n0];
  g.lg0 = __log2f((float)g.first);
  g.scale = (float)n0 / (__log2f((float)g.last) - g.lg0);
  return g;
}

// cnt(x) = #{k in [0, n0] : e0(k) <= x}: the log-spaced guess, exact
// correction by at most two steps, else bisection (any ascending grid).
template <class E>
__device__ __forceinline__ int count_le(E e0, int n0, double x, const GridGuess &g) {
  if (!(x >= g.first)) return 0;
  if (x >= g.last) return n0 + 1;
  int k = (int)((__log2f((float)x) - g.lg0) * g.scale);
  k = min(max(k, 0), n0 - 1);
#pragma unroll
  for (int it = 0; it < 3; it++) {
    if (e0(k) > x) {
      if (k == 0) return 0;  // cannot happen: x >= e0[0]
      k--;
    } else if (e0(k + 1) <= x) {
      k++;
    } else {
      return k + 1;
    }
  }
  int lo = 0, hi = n0 + 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (e0(mid) <= x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Guessed bracket of x (the first guess of count_le), clamped to [0, n0 - 1].
__device__ __forceinline__ int guess_index(double x, const GridGuess &g, int n0) {
  const int k = (int)((__log2f((float)x) - g.lg0) * g.scale);
  return min(max(k, 0), n0 - 1);
}

// count_le when the edges a = e0(k), b = e0(k + 1) around the guess k are
// already loaded: the exact search only when the guess misses.
template <class E>
__device__ __forceinline__ int count_le_at(E e0, int n0, double x, const GridGuess &g, int k, double a, double b) {
  if (!(x >= g.first)) return 0;
  if (x >= g.last) return n0 + 1;
  if (a <= x && x < b) return k + 1;
  return count_le(e0, n0, x, g);
}

template <int CTRL>
__device__ inline int dpp_int(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ inline int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ inline double dpp_d(double v) {
  const int lo = dpp_i<CTRL>(__double2loint(v));
  const int hi = dpp_i<CTRL>(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

__device__ inline double readlane_d(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

// Sum over the 64 lanes in a fixed order (pairs, quads, half rows, rows, then
// the four rows); every lane must be active.
__device__ inline double wave_sum(double v) {
  v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_d<0x141>(v);  // row_half_mirror
  v += dpp_d<0x140>(v);  // row_mirror
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

// min / max over the 64 lanes (quad, half-row and row exchanges, then the rows);
// every lane must be active
__device__ inline int wave_min_i(int v) {
  v = min(v, dpp_int<0xB1>(v));
  v = min(v, dpp_int<0x4E>(v));
  v = min(v, dpp_int<0x141>(v));
  v = min(v, dpp_int<0x140>(v));
  return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

__device__ inline int wave_max_i(int v) {
  v = max(v, dpp_int<0xB1>(v));
  v = max(v, dpp_int<0x4E>(v));
  v = max(v, dpp_int<0x141>(v));
  v = max(v, dpp_int<0x140>(v));
  return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// block-wide maximum
__device__ inline double block_max(double v, double *red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] = fmax(red[tid], red[tid + s]);
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

// 1/x to within an ulp or two: hardware estimate and two Newton steps
// (cheaper than the IEEE division sequence; the rebin weights need ~1e-15).
__device__ __forceinline__ double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// _rebin_spectrum (src/relutility.c:549-601), value of one output bin
// [elo_, ehi_] from the counts clo = cnt(elo_), chi = cnt(ehi_) of input
// edges at or below each end, cnt(x) = #{k : e0(k) <= x}: the reference's
// sequential search gives imin = max(clo-1, 0), imax = max(min(chi, n0)-1, 0);
// an overlap of zero width at the top edge gives 0.
template <class E, class F>
__device__ __forceinline__ double rebin_bin(double elo_, double ehi_, int clo, int chi, E e0, int n0, F f0) {
  const int imin = clo > 0 ? clo - 1 : 0;
  if (!((e0(0) <= ehi_) && (e0(n0) >= elo_)) || imin >= n0) return 0.0;
  int imax = chi < n0 ? chi : n0;
  if (imax > 0) imax--;
  double elo = elo_, ehi = ehi_;
  const double a0 = e0(imin), a1 = e0(imin + 1);
  if (elo < a0) elo = a0;
  const double b0 = e0(imax), b1 = e0(imax + 1);
  if (ehi > b1) ehi = b1;
  const double ra = rcp64(a1 - a0);
  if (imax == imin) return (ehi - elo) * ra * f0(imin);
  const double dmin = (a1 - elo) * ra;
  const double dmax = (ehi - b0) * rcp64(b1 - b0);
  double v = 0.0;
  v += f0(imin) * dmin + f0(imax) * dmax;
  for (int jj = imin + 1; jj <= imax - 1; jj++) v += f0(jj);
  return v;
}

// A fine radius' lerp factor between the node radii K+1 (lo) and K (hi) and
// its g window, lerped between the nodes' g_min and g_max (src/Relprofile.cpp:
// 249-300).  k_setup, k_integrate and k_fixup all take them from here (k_setup
// from its LDS copy of the node arrays, the others from the request's node
// arrays in global memory), so that the three agree bit for bit.
struct RadGeo {
  double ifac_r, gmin, gmax;
};
__device__ inline RadGeo rad_geo(const double *tre, const double *tgmin, const double *tgmax, double re, int K) {
#pragma clang fp contract(off)
  RadGeo o;
  o.ifac_r = (re - tre[K + 1]) / (tre[K] - tre[K + 1]);
  o.gmin = lerp1(o.ifac_r, tgmin[K + 1], tgmin[K]);
  o.gmax = lerp1(o.ifac_r, tgmax[K + 1], tgmax[K]);
  return o;
}

// block-wide deterministic sum (fixed tree order)
__device__ inline double block_sum(double v, double *red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

}  // namespace

// ---------------------------------------------------------------------------
// setup
// ---------------------------------------------------------------------------
// NZ: the most zones of a request in the batch (LP_ZONES, or MAX_ZONES with
// an ionisation gradient), which sizes the zone arrays in LDS
template <int NZ>
__global__ __launch_bounds__(SETUP_THREADS, 3) void k_setup(BatchDev b) {
  const int req = blockIdx.x;
  const int tid = threadIdx.x;
  const ReqParams rq = b.req[req];
  const DevRelTable &T = b.rt;
  const int nr = b.nr_fine;
  const int n_r = T.n_r, ng = T.n_g;
  if (rq.pre_status != 0) return;  // status already set by the host
  if (rq.model == MOD_XILLVER) {
    // the xillver models (LocalModel::xillver_model, src/LocalModel.cpp:
    // 104-128) have no relativistic part: the zone record carries the Ecut
    // (kTe) bracket, the angle distribution the inclination's bracket, so
    // that k_xill's sum over angles is the interpolation in the inclination
    // (interp_5d_tab, src/xilltable.c:878-995); k_primary rebins and adds the
    // primary
    if (tid == 0) {
      // the bracket through the request in global memory: a runtime index
      // into the local copy rq would move that copy to scratch
      const ReqParams &gq = b.req[req];
      const DevXillTable &xt = b.xt[rq.prim];
      ZoneRec o;
      o.rmean = 0.0;
      o.del_emit = 0.0;
      o.shift = 1.0;
      o.ncf = 1.0;
      o.ect_ind = xt.ect_dim >= 0 ? gq.xind[xt.ect_dim] : 0;
      o.ect_fac = xt.ect_dim >= 0 ? gq.xfac[xt.ect_dim] : 0.0;
      o.lxi_ind = xt.lxi_dim >= 0 ? gq.xind[xt.lxi_dim] : 0;
      o.lxi_fac = xt.lxi_dim >= 0 ? gq.xfac[xt.lxi_dim] : 0.0;
      b.zone[rq.zone_off] = o;
    }
    double *d = b.dist + rq.dist_off;
    for (int a = tid; a < rq.n_incl; a += SETUP_THREADS)
      d[a] = (a == rq.xinc_ind) ? 1.0 - rq.xinc_fac : ((a == rq.xinc_ind + 1) ? rq.xinc_fac : 0.0);
    return;
  }
  if (rq.model == MOD_RELXILL || rq.model == MOD_RELXILLLP) {
    // the zones' convolutions add into the output (k_conv)
    double *out = b.out + rq.flux_off;
    for (int j = tid; j < rq.n_in; j += SETUP_THREADS) out[j] = 0.0;
  }

  __shared__ double s_tre[MAX_NR_TAB], s_tgmin[MAX_NR_TAB], s_tgmax[MAX_NR_TAB];
  __shared__ double s_jr[MAX_LP_R], s_je[MAX_LP_R], s_jd[MAX_LP_R], s_ji[MAX_LP_R];
  __shared__ double s_red[SETUP_THREADS];
  __shared__ double s_dist[NZ * MAX_INCL];
  __shared__ int s_st, s_ia, s_im, s_rmin;
  __shared__ double s_fa, s_fm, s_rms;
  constexpr int MAX_RC = MAX_NR_FINE / RAD_CHUNK;
  // energy window of each (radius chunk, zone): bins [s_ulo, s_uhi]
  __shared__ int s_ulo[MAX_RC * NZ], s_uhi[MAX_RC * NZ];
  __shared__ int s_pref[MAX_RC * NZ + 1], s_prefs[MAX_RC * NZ + 1];
  __shared__ int s_base, s_bases;
  const int n_rc = (b.nr_fine + RAD_CHUNK - 1) / RAD_CHUNK;
  const int nz = rq.n_zones;
  const double *zg = b.req[req].zgrid;
  for (int c = tid; c < n_rc * NZ; c += SETUP_THREADS) {
    s_ulo[c] = 0x7fffffff;
    s_uhi[c] = -1;
  }

  // profile grid: caller grid shifted to line-energy units (XspecSpectrum.h:61-76)
  const double *E;
  if (rq.pgrid_off >= 0) {
    double *pg = b.pgrid + rq.pgrid_off;
    const double *ein = b.ener_in + rq.ener_off;
    for (int k = tid; k <= rq.n_in; k += SETUP_THREADS) {
      double e = ein[k];
      if (rq.z > 0) e *= (1 + rq.z);
      pg[k] = e / rq.lineE;
    }
    E = pg;
  } else {
    E = b.cg.ener;
  }

  if (tid == 0) {
    int st = 0;
    const double rms = rq.rms;
    if (!(rq.rout > rq.rin) || !(rq.rin >= rms)) st = 10;
    const double mu0 = cos(rq.incl);
    const int ia = bsearch_asc(T.a, T.n_a, (float)rq.a);
    const int im = bsearch_asc(T.mu0, T.n_mu0, (float)mu0);
    const float fa = ((float)rq.a - T.a[ia]) / (T.a[ia + 1] - T.a[ia]);
    const float fm = ((float)mu0 - T.mu0[im]) / (T.mu0[im + 1] - T.mu0[im]);
    if (!(fm >= 0 && fm <= 1 && fa >= 0 && fa <= 1)) st = 11;
    s_st = st;
    s_ia = ia;
    s_im = im;
    s_fa = fa;
    s_fm = fm;
    s_rms = rms;
  }
  __syncthreads();
  if (s_st) {
    if (tid == 0) b.status[req] = s_st;
    return;
  }
  const int ia = s_ia, im = s_im;
  const double fa = s_fa, fm = s_fm;
  const size_t n00 = (size_t)ia * T.n_mu0 + im, n10 = (size_t)(ia + 1) * T.n_mu0 + im;
  const size_t n01 = (size_t)ia * T.n_mu0 + im + 1, n11 = (size_t)(ia + 1) * T.n_mu0 + im + 1;

  // node radii: lerp in a only (src/Relprofile.cpp:190-193)
  for (int ir = tid; ir < n_r; ir += SETUP_THREADS) {
    s_tre[ir] = lerp1(fa, (double)T.r[n00 * n_r + ir], (double)T.r[n10 * n_r + ir]);
  }
  __syncthreads();
  if (tid == 0) {
    int st = 0;
    const double rms = s_rms;
    // ISCO snap (:195-200)
    if ((s_tre[n_r - 1] > rms) && ((s_tre[n_r - 1] - rms) / s_tre[n_r - 1] < 1e-3)) {
      s_tre[n_r - 1] = rms;
    }
    const int ind_rmin = bsearch_desc(s_tre, n_r, rq.rin);
    const int ind_rmax = bsearch_desc(s_tre, n_r, rq.rout);
    // 1000 rg snap (:235-238)
    if (s_tre[ind_rmax] < 1000.0 && s_tre[ind_rmax] * 1.01 > 1000.0) s_tre[ind_rmax] = 1000.0;
    if (!(ind_rmin > 0) || !(s_tre[ind_rmin + 1] <= rq.rin) || !(s_tre[ind_rmin] >= rq.rin) ||
        !(s_tre[ind_rmax + 1] <= rq.rout) || !(s_tre[ind_rmax] >= rq.rout) ||
        !(ind_rmax <= ind_rmin) || !(rq.rout <= 1000.0)) {
      st = 12;
    }
    s_rmin = ind_rmin;
    s_st = st;
  }
  // node values: bilinear in (a, mu0) (interpol_a_mu0, :39-80)
  for (int ir = tid; ir < n_r; ir += SETUP_THREADS) {
    s_tgmin[ir] = lerp2f(fa, fm, T.gmin[n00 * n_r + ir], T.gmin[n10 * n_r + ir],
                         T.gmin[n01 * n_r + ir], T.gmin[n11 * n_r + ir]);
    s_tgmax[ir] = lerp2f(fa, fm, T.gmax[n00 * n_r + ir], T.gmax[n10 * n_r + ir],
                         T.gmax[n01 * n_r + ir], T.gmax[n11 * n_r + ir]);
  }
  if (rq.n_incl > 0) {  // node rows for the angle distribution below
    double *node = b.node + rq.node_off;
    const size_t nq = (size_t)n_r * ng;
    const size_t s00 = n00 * nq, s10 = n10 * nq, s01 = n01 * nq, s11 = n11 * nq;
    for (size_t q = tid; q < nq; q += SETUP_THREADS) {
      node[q * 2 + 0] = lerp2f(fa, fm, T.trff1[s00 + q], T.trff1[s10 + q], T.trff1[s01 + q], T.trff1[s11 + q]);
      node[q * 2 + 1] = lerp2f(fa, fm, T.trff2[s00 + q], T.trff2[s10 + q], T.trff2[s01 + q], T.trff2[s11 + q]);
      if (rq.need_cosne) {
        node[nq * 2 + q * 2 + 0] = lerp2f(fa, fm, T.cosne1[s00 + q], T.cosne1[s10 + q], T.cosne1[s01 + q], T.cosne1[s11 + q]);
        node[nq * 2 + q * 2 + 1] = lerp2f(fa, fm, T.cosne2[s00 + q], T.cosne2[s10 + q], T.cosne2[s01 + q], T.cosne2[s11 + q]);
      }
    }
  }
  // lamp-post table node (interpol_lptable, src/Rellp.cpp:219-247)
  const bool lp = (rq.emis_type == EMIS_LP);
  const DevLpTable &L = b.lp;
  __syncthreads();
  if (s_st) {
    if (tid == 0) b.status[req] = s_st;
    return;
  }
  {  // the node arrays the radius records are rebuilt from (rad_geo)
    double *gn = b.rnode + (size_t)req * RAD_NODE_ARRAYS * n_r;
    for (int ir = tid; ir < n_r; ir += SETUP_THREADS) {
      gn[ir] = s_tre[ir];
      gn[n_r + ir] = s_tgmin[ir];
      gn[2 * n_r + ir] = s_tgmax[ir];
    }
  }
  const int ntr = L.n_r;
  if (lp) lp_rows(L, rq, s_jr, s_je, s_jd, s_ji, tid, SETUP_THREADS);

  // fine radii: node -> fine map (:249-300), emissivity, energy window (:859-878)
  RadRec *rad = b.rad + (size_t)req * nr;
  const GridGuess gE = grid_guess(E, rq.n_e);
  auto ge = [&](int k) { return E[k]; };
  double area = 0.0;
  double cmax = 0.0;  // largest weight of a radius with a window
  int err = 0;
  // lamp-post emissivity of a fine radius with the flux boost (src/Rellp.cpp:249-253,
  // src/Relphysics.cpp:301-311)
  auto lp_emis = [&](double re) {
    LpFine f = lp_on_fine(re, s_jr, s_je, s_jd, s_ji, ntr);
    if (f.err) err = f.err;
    double boost = pow(gi_potential_lp(re, rq.a, rq.height, rq.beta, f.del_emit), rq.gamma);
    if (rq.beta > 1e-6) {
      const double d = sqrt(1.0 - rq.beta * rq.beta) / (1.0 + rq.beta * cos(f.del_emit));
      boost *= d * d;
    }
    return f.emis * boost;
  };
  // returning radiation (add_returnrad_emis, src/Rellp.cpp:483-507, with
  // get_rrad_emis_corona, src/Relreturn_Corona.cpp:291-321): the direct
  // emissivity of every fine radius (parked in the radius records), at the
  // table zones' centres (inv_rebin_mean), the zones' returning emissivity
  // sum_re frac_g g^(Gamma-1) tf_r emis(re) (calc_rrad_emis_corona, :119-151),
  // log-interpolated back onto the fine radii below
  __shared__ double s_rr_rad[MAX_RR], s_rr_in[MAX_RR], s_rr_ret[MAX_RR], s_rr_corr[2];
  __shared__ double s_rr_cf[MAX_RR], s_rr_cg[MAX_RR];
  const bool rrad = lp && rq.return_rad != 0 && s_st == 0;
  const int rr_n = rq.rr_n;
  if (rrad) {
    for (int ii = tid; ii < nr; ii += SETUP_THREADS) rad[ii].coef = lp_emis(fine_radius(rq.rin, rq.rout, ii, nr));
    const DevRradTable &R = b.rr;
    const int nt = R.n_rad, k = rq.rr_k, klo = rq.rr_klo;
    const double *rlo_t = R.rlo + (size_t)k * nt, *rhi_t = R.rhi + (size_t)k * nt;
    const double rin_f = fine_radius(rq.rin, rq.rout, nr - 1, nr), rout_f = fine_radius(rq.rin, rq.rout, 0, nr);
    for (int i = tid; i < rr_n; i += SETUP_THREADS) {
      const double lo = (i == 0) ? rin_f : rlo_t[klo + i];
      const double hi = (i == rr_n - 1) ? rout_f : rhi_t[klo + i];
      const double rm = 0.5 * (lo + hi);
      s_rr_rad[i] = rm;
      // relxilllp above spin 0: the correction factors of the relxill zone
      // the table zone's centre falls in (rebin_corrfactors_to_rradtable_grid,
      // src/Relreturn_Corona.cpp:248-289; a centre beyond the last edge takes
      // the last zone)
      double cf = 1.0, cg = 1.0;
      if (rq.rr_corr) {
        const int n0 = nz;
        int ind = bsearch_asc(zg, n0 + 1, rm);
        if (rm < zg[0]) ind = 0;
        else if (rm > zg[n0] || ind > n0 - 1) ind = n0 - 1;
        cf = b.rrcorr[rq.rr_corr_off + 2 * ind];
        cg = b.rrcorr[rq.rr_corr_off + 2 * ind + 1];
      }
      s_rr_cf[i] = cf;
      s_rr_cg[i] = cg;
    }
    if (tid < 2) {  // ring-area corrections of the end zones (src/Relreturn_Table.cpp:415-435)
      const int i = (tid == 0) ? 0 : rr_n - 1;
      double rl = rlo_t[klo + i];
      if (klo + i == 0 && rl > rq.rms) rl = rq.rms;
      const double rh = rhi_t[klo + i];
      const double lo = (i == 0) ? rin_f : rlo_t[klo + i];
      const double hi = (i == rr_n - 1) ? rout_f : rhi_t[klo + i];
      s_rr_corr[tid] = (0.5 * (lo + hi) * (hi - lo)) / (0.5 * (rl + rh) * (rh - rl));
    }
    __syncthreads();
    for (int i = tid; i < rr_n; i += SETUP_THREADS) {
      // inv_rebin_mean (src/relutility.c:636-663): the fine radii re(ii) > rad >= re(ii+1)
      const double x = s_rr_rad[i];
      int lo = 0, hi = nr - 1;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (fine_radius(rq.rin, rq.rout, mid, nr) > x) lo = mid; else hi = mid;
      }
      const double r0 = fine_radius(rq.rin, rq.rout, lo, nr), r1 = fine_radius(rq.rin, rq.rout, lo + 1, nr);
      s_rr_in[i] = lerp1((x - r1) / (r0 - r1), rad[lo + 1].coef, rad[lo].coef);
    }
    __syncthreads();
    // calc_rrad_emis_corona (src/Relreturn_Corona.cpp:119-151): one wavefront
    // per incident zone i, lane j its emitting zones (at most 64), summed in a
    // fixed order; the frac_g weight g^(Gamma-1), or with a g-shift factor off
    // 1 by more than 1e-3 of the emitting zone its corrected boost over g
    // (calc_rrad_emis_zone, :87-108); the sum times the incident zone's flux
    // factor
    const int ng = R.n_g;
    const int wv = tid >> 6, lane = tid & 63;
    for (int i = wv; i < rr_n; i += SETUP_THREADS / 64) {
      double term = 0.0;
      const int j = lane;
      if (j < rr_n) {
        const size_t t = (size_t)k * nt * nt + (size_t)(klo + i) * nt + (klo + j);
        const double gmin = R.gmin[t], gmax = R.gmax[t];
        const double cgj = s_rr_cg[j];
        double ez = 0.0;
        for (int jj = 0; jj < ng; jj++) {
          const double g = (jj + 0.5) / ng * (gmax - gmin) + gmin;  // get_gfac_grid (src/Relreturn_Datastruct.cpp:51-59)
          double v = R.frac_g[t * ng + jj];
          if (fabs(cgj - 1) > 1e-3) v *= gshift_boost(cgj, g, rq.gamma) / g;
          else if (fabs(g - 1) > 1e-3) v *= pow(g, rq.gamma - 1);
          ez += v;
        }
        double tf = R.tf_r[t];
        if (j == 0 && i < rr_n - 1) tf *= s_rr_corr[0];
        if (j == rr_n - 1 && i < rr_n - 2) tf *= s_rr_corr[1];
        term = ez * tf * s_rr_in[j];
      }
      const double sum = wave_sum(term);
      if (lane == 0) s_rr_ret[i] = sum * s_rr_cf[i];
    }
    __syncthreads();
  }
  for (int ii0 = 0; ii0 < nr; ii0 += SETUP_THREADS) {  // same trip count in every wavefront
    const int ii = ii0 + tid;
    int wlo = 0x7fffffff, whi = -1;
    int izone = -1;
    if (ii < nr) {
    const double re = fine_radius(rq.rin, rq.rout, ii, nr);
    const double rprev = (ii > 0) ? fine_radius(rq.rin, rq.rout, ii - 1, nr) : 0.0;
    const double rnext = (ii < nr - 1) ? fine_radius(rq.rin, rq.rout, ii + 1, nr) : 0.0;
    // K = largest k <= ind_rmin with tre[k] > re: the walk of :250-265 starts
    // at the Rin bracket and only moves outwards, so a fine radius rounded
    // just below the innermost node extrapolates that bracket (ifac_r < 0)
    int lo = 0, hi = n_r;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (s_tre[mid] > re) lo = mid + 1; else hi = mid;
    }
    int K = min(lo - 1, s_rmin);
    if (K < 0) {  // (re beyond the outermost node: the request fails, K stays in bounds)
      if (!(re - 1000.0 <= 1e-6)) err = 13;
      K = 0;
    }
    if (K >= n_r - 1) {  // cannot happen for rin >= tre[n_r-1]; keep in bounds
      K = n_r - 2;
      err = 13;
    }
    const RadGeo geo = rad_geo(s_tre, s_tgmin, s_tgmax, re, K);
    const double ifac_r = geo.ifac_r;
    if (ifac_r > 1.0 && K > 0) err = 14;
    const double gmin = geo.gmin;
    const double gmax = geo.gmax;
    const double tz = trapez_half(rprev, re, rnext, ii, nr);
    const double weight = tz / 2;
    double emis;
    if (!lp) {  // broken power law (get_emis_bkn, src/Rellp.cpp:421-437)
      const double alpha = (re > rq.rbr) ? rq.emis2 : rq.emis1;
      emis = pow(re / rq.rbr, -alpha);
      area += emis * (tz * 2);
    } else if (!rrad) {
      emis = lp_emis(re);
    } else {
      // the returning emissivity at this radius (rebin_emisprofile_on_radial_grid,
      // src/Rellp.cpp:120-175: linear radius factor, log-interpolated values)
      const int kk = min(max(bsearch_asc(s_rr_rad, rr_n, re), 0), rr_n - 2);
      const double fr = (re - s_rr_rad[kk]) / (s_rr_rad[kk + 1] - s_rr_rad[kk]);
      const double er = exp(fr * log(s_rr_ret[kk + 1]) + (1.0 - fr) * log(s_rr_ret[kk]));
      emis = (rq.return_rad == 1) ? rad[ii].coef + er : er;
    }
    // energy window
    double egmin = gmin, egmax = gmax;
    int ielo = 1, iehi = 0;
    if ((egmax > E[0]) && (egmin < E[rq.n_e])) {
      if (egmin < E[0]) egmin = E[0];
      if (egmax > E[rq.n_e]) egmax = E[rq.n_e];
      // binary_search (src/relutility.c:155-171) on [E0, E_n]: count - 1, at most n-1
      ielo = min(count_le(ge, rq.n_e, egmin, gE) - 1, rq.n_e - 1);
      iehi = min(count_le(ge, rq.n_e, egmax, gE) - 1, rq.n_e - 1);
    }
    // zone of the radius (binary_search on the zone edges, src/Relprofile.cpp:874)
    izone = (nz > 1) ? bsearch_asc(zg, nz + 1, re) : 0;
    const double coef = emis / (gmax - gmin) * weight;  // times ReqNorm.fac
    rad[ii] = rad_pack(coef, K, izone, ielo, iehi);
    if (ielo <= iehi) {
      wlo = ielo;
      whi = iehi;
      cmax = fmax(cmax, fabs(coef));
    }
    }  // ii < nr
    // energy window of the radius chunk (this wavefront's 64 consecutive
    // radii), per zone present in the chunk
    const int zlo = wave_min_i(ii < nr ? izone : 0x7fffffff);
    const int zhi = wave_max_i(ii < nr ? izone : -1);
    for (int z = zlo; z <= zhi; z++) {
      const int zl = wave_min_i(izone == z ? wlo : 0x7fffffff);
      const int zh = wave_max_i(izone == z ? whi : -1);
      if ((tid & 63) == 0) {
        s_ulo[(ii / RAD_CHUNK) * NZ + z] = zl;
        s_uhi[(ii / RAD_CHUNK) * NZ + z] = zh;
      }
    }
  }
  if (err) atomicMax(&s_st, err);
  // BKN normalisation: sum emis * 2 pi r dr (norm_emis_profile, src/Rellp.cpp:91-109),
  // applied through ReqNorm instead of a second pass over the records
  const double total = block_sum(area, s_red);
  const double cmax_all = block_max(cmax, s_red);
  if (s_st) {
    if (tid == 0) b.status[req] = s_st;
    return;
  }
  const double fac = lp ? 1.0 : 1.0 / total;
  if (tid == 0) {
    ReqNorm n;
    n.fac = fac;
    n.fa = fa;
    n.fm = fm;
    n.ia = ia;
    n.im = im;
    n.r1 = fine_radius_r1(rq.rout);
    n.r2 = fine_radius_r1(rq.rin);
    const double m = cmax_all * fac;
    int k = 0;
    if (m > 0.0 && m < INFINITY) k = min(1000, max(-1000, -ilogb(m)));
    n.cmul = ldexp(fac, k);
    n.cinv = ldexp(1.0, -k);
    b.rnorm[req] = n;
  }

  // reflection fractions (calc_refl_frac, src/Rellp.cpp:36-89)
  if (lp && tid == 0) {
    auto fr = [&](int ii) { return fine_radius(rq.rin, rq.rout, ii, nr); };
    int klo = 0, khi = nr - 1;
    while (khi - klo > 1) {
      int k = (khi + klo) / 2;
      if (fr(k) < rq.rin) khi = k; else klo = k;
    }
    const int ibh = klo;
    klo = 0;
    khi = nr - 1;
    while (khi - klo > 1) {
      int k = (khi + klo) / 2;
      if (fr(k) < rq.rout) khi = k; else klo = k;
    }
    const int iad = klo;
    double del_bh = lp_on_fine(fr(ibh), s_jr, s_je, s_jd, s_ji, ntr).del_emit;
    double del_ad = lp_on_fine(fr(iad), s_jr, s_je, s_jd, s_ji, ntr).del_emit;
    double dmax = s_jd[ntr - 1];
    if (dmax < PI / 2.0) dmax = PI / 2.0;
    if (rq.beta > 1e-6) {
      del_bh = relat_abberation(del_bh, -1. * rq.beta);
      del_ad = relat_abberation(del_ad, -1. * rq.beta);
    }
    LpRefl o;
    o.f_bh = 0.5 * (1.0 - cos(del_bh));
    o.f_ad = 0.5 * (cos(del_bh) - cos(del_ad));
    o.f_inf_rest = 0.5 * (1.0 + cos(dmax));
    o.f_inf = (rq.beta > 1e-6) ? 0.5 * (1.0 + cos(relat_abberation(dmax, -1. * rq.beta))) : o.f_inf_rest;
    o.refl_frac = o.f_ad / o.f_inf;
    if (!(o.f_inf + o.f_ad < 1.0)) s_st = 20;
    b.refl[req] = o;
  }

  // zones of a relxill request (IonGradient, src/IonGradient.cpp:182-209,
  // :421-430): the centre radius, its emission angle (inv_rebin_mean of the
  // fine-grid angles, src/relutility.c:636-663), the source -> disk energy
  // shift, the Ecut the zone sees with its xillver bracket (float factor,
  // clamps, ensure_ecut_within_boundarys, src/xilltable.c:1054-1125) and the
  // normalisation change of the incident spectrum (src/Xillspec.cpp:408-435)
  const bool relxill_model = (rq.model == MOD_RELXILL || rq.model == MOD_RELXILLLP);
  if (relxill_model && s_st == 0) {
    ZoneRec *zr = b.zone + rq.zone_off;
    const DevXillTable &xt = b.xt[rq.prim];
    __shared__ double s_zect[NZ];
    for (int z = tid; z < nz; z += SETUP_THREADS) {
      double ect;
      zr[z] = zone_record(xt, b.req[req], z, lp, s_jr, s_je, s_jd, s_ji, ntr, nr, &ect);
      s_zect[z] = ect;
    }
    __syncthreads();
    if (lp && rq.prim == PRIM_NTHCOMP) {
      // the nthcomp norms of the source and of every zone's kTe (k_nth)
      const NthRes *nr = b.nth + rq.nth_off;
      for (int z = tid; z < nz; z += SETUP_THREADS) zr[z].ncf = nr[1 + z].nf / nr[0].nf;
    } else if (lp) {
      const double src = 1. / (block_sum(xill_norm_part(b.cg, rq.xect_src, rq.xgam, tid, SETUP_THREADS),
                                         s_red) / NORM_XILLVER_TABLE);
      for (int z = 0; z < nz; z++) {
        const double ez = s_zect[z];
        const double disk =
            1. / (block_sum(xill_norm_part(b.cg, ez, rq.xgam, tid, SETUP_THREADS), s_red) / NORM_XILLVER_TABLE);
        if (tid == 0) zr[z].ncf = disk / src;
      }
    }
    __syncthreads();
  }

  // emission-angle distribution (:908-938).  Lane = fine radius, loop over
  // the 2 n_g (g*, branch) columns: a radius' record is read once and its
  // node rows are contiguous.  Neighbouring radii mostly fall into the same
  // angle bin, so a wavefront whose valid lanes share the bin adds one wave
  // sum to it; otherwise every lane adds its own term.
  if (rq.n_incl > 0) {
    const int nd = rq.n_incl * nz;  // [zone][angle bin]
    for (int k = tid; k < nd; k += SETUP_THREADS) s_dist[k] = 0.0;
    __syncthreads();
    const double *node = b.node + rq.node_off;
    const size_t nq2 = (size_t)n_r * ng * 2;
    const int cols = ng * 2;
    const double step = (1.0 - 2 * GFAC_H) / (ng - 1);
    const double gs0 = gstar_node(step, 0), gs1 = gstar_node(step, 1);
    // per g*: the node value and d_gstar / sqrt(g* (1 - g*)) (s_red is free here)
    static_assert(2 * MAX_NG <= SETUP_THREADS, "g* factors must fit the reduction buffer");
    double *s_gs = s_red, *s_gw = s_red + MAX_NG;
    for (int g = tid; g < ng; g += SETUP_THREADS) {
      const double gs = gstar_node(step, g);
      const double dgs = (g == 0 || g == ng - 1) ? 0.5 * (gs1 - gs0) + GFAC_H : gs1 - gs0;
      s_gs[g] = gs;
      s_gw[g] = dgs / sqrt(gs - gs * gs);
    }
    __syncthreads();
    const int lane = tid & 63;
    int derr = 0;
    for (int i0 = 0; i0 < nr; i0 += SETUP_THREADS) {
      const int ii = i0 + tid;
      bool live = ii < nr;
      RadRec rec = rad_pack(0.0, 0, 0, 1, 0);
      if (live) {
        rec = rad[ii];
        live = rad_ielo(rec) <= rad_iehi(rec);
      }
      if (__ballot(live) == 0) continue;
      const double re = live ? fine_radius(rq.rin, rq.rout, ii, nr) : 0.0;
      const int K = rad_K(rec), rzone = rad_zone(rec);
      RadGeo r = {0.0, 0.0, 0.0};
      if (live) r = rad_geo(s_tre, s_tgmin, s_tgmax, re, K);
      const double emisw = rec.coef * (r.gmax - r.gmin) * fac;  // emis * weight
      const double *tlo = node + (size_t)(K + 1) * cols, *thi = node + (size_t)K * cols;
      for (int q = 0; q < cols; q++) {
        const int g = q >> 1;
        const double gs = s_gs[g];
        double tmp = 0.0;
        int imu = -1;
        bool ok = false;
        if (live) {
          const double trff = lerp1(r.ifac_r, tlo[q], thi[q]);
          const double cosne = lerp1(r.ifac_r, tlo[nq2 + q], thi[nq2 + q]);
          const double gv = gs * (r.gmax - r.gmin) + r.gmin;
          imu = ((int)(rq.n_incl * (1 - cosne) + 1)) - 1;
          const double pr = 2 * PI * gv * re;
          tmp = re * (pr * pr) * trff * emisw * s_gw[g];
          if (imu < 0 || imu >= rq.n_incl) derr = 40;
          else if (tmp != tmp) derr = 41;
          else ok = true;
          imu += rzone * rq.n_incl;  // the radius' zone
        }
        const unsigned long long m_ok = __ballot(ok);
        if (m_ok == 0) continue;
        const int first = __ffsll((long long)m_ok) - 1;
        const int k0 = __builtin_amdgcn_readlane(imu, first);
        if (__ballot(ok && imu != k0) == 0) {
          const double sum = wave_sum(ok ? tmp : 0.0);
          if (lane == first) atomicAdd(&s_dist[k0], sum);
        } else if (ok) {
          atomicAdd(&s_dist[imu], tmp);
        }
      }
    }
    if (derr) atomicMax(&s_st, derr);
    __syncthreads();
    for (int k = tid; k < nd; k += SETUP_THREADS) b.dist[rq.dist_off + k] = s_dist[k];
  }
  __syncthreads();
  if (s_st == 0) {
    // the profile accumulates with atomic adds: clear the bins any window
    // reaches (k_finalize writes zeros elsewhere)
    int jlo = 0x7fffffff, jhi = -1;
    for (int c = 0; c < n_rc; c++) {
      for (int z = 0; z < nz; z++) {
        jlo = min(jlo, s_ulo[c * NZ + z]);
        jhi = max(jhi, s_uhi[c * NZ + z]);
      }
    }
    if (tid == 0) b.reach[req] = BinRange{jlo, jhi};
    double *acc = profile_acc(b, rq);
    // (no window reaches the grid: jlo = INT_MAX > jhi = -1, nothing to clear;
    // jlo + tid must not be formed then, it would overflow); every zone's
    // profile over the union of the windows
    if (jlo <= jhi)
      for (int z = 0; z < nz; z++)
        for (int j = jlo + tid; j <= jhi; j += SETUP_THREADS) acc[(size_t)z * rq.n_e + j] = 0.0;
    // work items: the window of each (radius chunk, zone) cut into segments
    // of seg_bins; the item's chunk and zone share its rc field.  An item's
    // time grows with its bins, so segments of at least S / 4 bins go to the
    // front of the list and the shorter last segments of windows to its back:
    // k_integrate hands out the long ones first and the wavefronts finish
    // together on short ones.
    const int S = b.seg_bins;
    const int thr = max(1, S / 4);
    const int n_cz = n_rc * nz;  // pairs q = c nz + z; their windows sit at c NZ + z
    if (tid == 0) {
      int tot = 0, tots = 0;
      for (int q = 0; q < n_cz; q++) {
        const int w = (q / nz) * NZ + q % nz;
        s_pref[q] = tot;
        s_prefs[q] = tots;
        if (s_uhi[w] >= s_ulo[w]) {
          const int len = s_uhi[w] - s_ulo[w] + 1;
          const int nseg = (len + S - 1) / S;
          const int nlong = nseg - 1 + (len - (nseg - 1) * S >= thr ? 1 : 0);
          tot += nlong;
          tots += nseg - nlong;
        }
      }
      int base = atomicAdd(&b.ctr[CTR_ITEMS], tot);
      const int bases = atomicAdd(&b.ctr[CTR_ITEMS_S], tots);
      // cannot happen: the host sizes the list for every bin of every request
      if (base + tot > b.item_cap || bases + tots > b.item_cap) {
        atomicOr(&b.ctr[CTR_ITEMS_OVF], 1);
        s_st = ST_WORKLIST;
        base = -1;
      }
      s_base = base;
      s_bases = bases;
    }
    __syncthreads();
    if (s_base >= 0) {
      for (int q = tid; q < n_cz; q += SETUP_THREADS) {
        const int c = q / nz, z = q % nz, w = c * NZ + z;
        int k = s_base + s_pref[q], ks = s_bases + s_prefs[q];
        for (int j0 = s_ulo[w]; j0 <= s_uhi[w]; j0 += S) {
          WorkItem it;
          it.req = req;
          it.rc = c | (z << 16);
          it.j0 = j0;
          it.j1 = min(s_uhi[w] + 1, j0 + S);
          if (it.j1 - it.j0 >= thr) b.items[k++] = it;
          else b.items[b.item_cap - 1 - ks++] = it;
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) b.status[req] = s_st;
}

// relxilllp's returning radiation above spin 0: the xillver correction
// factors of one zone (calc_rrad_corr_factors, src/Relxill.cpp:164-182, with
// get_xillver_fluxcorrection_factors, src/Xillspec.cpp:500-526).  The zone's
// xillver spectra for every inclination (16 table corners with the zone's own
// Ecut bracket, as k_xill reads them) are averaged with the weights of
// calc_angle_averaged_xill_spec (:109-126: 0.5 cos(i) / n_incl with the
// table's inclination, in degrees, read as radians), then
//   flux factor = energy flux of the average in 0.1-1000 keV over that of the
//                 normalised cut-off power law on the 500-bin grid (:344-362)
//   g factor    = photon flux of the average in 0.15-500 keV over that of the
//                 average shifted 1.5 times lower in energy (rebinned, times
//                 2/3), over 1.5^Gamma (:454-487).
// One workgroup per (request, zone); the average sits in dynamic LDS; the
// sums are fixed-order block sums.  k_setup reads the factors (it runs next).
constexpr int RRC_THREADS = 256;
__global__ __launch_bounds__(RRC_THREADS) void k_rrad_corr(BatchDev b, const XillWork *__restrict__ work) {
  extern __shared__ double s_avg[];
  __shared__ double s_jr[MAX_LP_R], s_je[MAX_LP_R], s_jd[MAX_LP_R], s_ji[MAX_LP_R];
  __shared__ double s_red[RRC_THREADS];
  const XillWork wk = work[blockIdx.x];
  const int req = wk.req, z = wk.zone, tid = threadIdx.x;
  const ReqParams &rq = b.req[req];
  if (rq.pre_status != 0) return;  // uniform over the block
  const DevXillTable xt = b.xt[rq.prim];
  const double dd = rq.xdens_div;
  lp_rows(b.lp, rq, s_jr, s_je, s_jd, s_ji, tid, RRC_THREADS);
  double ect;
  const ZoneRec zr = zone_record(xt, rq, z, true, s_jr, s_je, s_jd, s_ji, b.lp.n_r, b.nr_fine, &ect);
  int i0[4];
  double f[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {  // the zone's own Ecut and logXi brackets
    const bool e = (k == xt.ect_dim), l = (k == xt.lxi_dim);
    i0[k] = e ? zr.ect_ind : (l ? zr.lxi_ind : rq.xind[k]);
    f[k] = e ? zr.ect_fac : (l ? zr.lxi_fac : rq.xfac[k]);
  }
  // the 16 corners in the reference's order (interp_5d_tab_incl, src/xilltable.c:812-876)
  constexpr unsigned char kCorner[16] = {0x0, 0x8, 0x4, 0x2, 0xC, 0xA, 0x6, 0xE,
                                         0x1, 0x9, 0x5, 0x3, 0xD, 0xB, 0x7, 0xF};
  double w[16];
  unsigned row[16];
#pragma unroll
  for (int c = 0; c < 16; c++) {
    int o[4];
    double fk[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      o[k] = (kCorner[c] >> (3 - k)) & 1;
      fk[k] = o[k] ? f[k] : (1.0 - f[k]);
    }
    w[c] = fk[0] * fk[1] * fk[2] * fk[3];
    const unsigned r = (((unsigned)(i0[0] + o[0]) * xt.n_val[1] + (i0[1] + o[1])) * xt.n_val[2] + (i0[2] + o[2])) *
                           xt.n_val[3] + (i0[3] + o[3]);
    row[c] = r * xt.n_incl;
  }
  const int ne = xt.n_ener, n_incl = xt.n_incl;
  const int n_o = xt.n_out > 1 ? 2 : 1;
  const double *E = xt.ener;
  for (int e = tid; e < ne; e += RRC_THREADS) {
    double acc = 0.0;
    for (int a = 0; a < n_incl; a++) {
      const double incl_deg = (double)xt.incl_vals[a] * 180 / PI;
      const double wa = 0.5 * cos(incl_deg * PI / 180) / n_incl;
      double sv = 0.0;
      for (int o = 0; o < n_o; o++) {  // (a six-parameter table: both brackets of its first axis)
        const double wo = n_o == 1 ? 1.0 : (o ? rq.xo_fac : 1.0 - rq.xo_fac);
        const size_t ob = (size_t)(rq.xo_ind + o) * xt.outer_rows;
        double so = 0.0;
#pragma unroll
        for (int c = 0; c < 16; c++) so += w[c] * xval(xt.spec[(ob + row[c] + a) * ne + e], dd);
        sv += wo * so;
      }
      acc += wa * sv;
    }
    s_avg[e] = acc;
  }
  __syncthreads();
  // band sums of the average, and of its shifted copy
  const double gref = 2. / 3.;
  const GridGuess gx = grid_guess(E, ne);
  auto ge = [&](int k) { return E[k]; };
  auto fav = [&](int k) { return s_avg[k]; };
  double ef = 0.0, pf = 0.0, pz = 0.0;
  for (int e = tid; e < ne; e += RRC_THREADS) {
    const double e0 = E[e], e1 = E[e + 1];
    if (e0 >= 0.1 && e1 <= 1000.0) ef += s_avg[e] * 0.5 * (e0 + e1);
    if (e0 >= 0.15 && e1 <= 500.0) {
      pf += s_avg[e];
      const double lo = e0 / gref, hi = e1 / gref;
      pz += rebin_bin(lo, hi, count_le(ge, ne, lo, gx), count_le(ge, ne, hi, gx), ge, ne, fav) * gref;
    }
  }
  const double s_ef = block_sum(ef, s_red), s_pf = block_sum(pf, s_red), s_pz = block_sum(pz, s_red);
  // the normalised primary of the zone (Ecut / kTe it sees, the source's
  // Gamma) on the 500-bin grid (calc_normalized_primary_spectrum, :302-332):
  // the cut-off power law here, nthcomp from its solve (k_nth)
  double s_pe;
  if (rq.prim == PRIM_NTHCOMP) {
    s_pe = b.nth[rq.nth_off + 1 + z].pe;
  } else {
    const double nf =
        1. / (block_sum(xill_norm_part(b.cg, ect, rq.xgam, tid, RRC_THREADS), s_red) / NORM_XILLVER_TABLE);
    double pe = 0.0;
    for (int k = tid; k < N_ENER_COARSE; k += RRC_THREADS) {
      const double e0 = b.cg.coarse[k], e1 = b.cg.coarse[k + 1];
      if (e0 >= 0.1 && e1 <= 1000.0)
        pe += cutoffpl_bin(e0, e1, b.cg.coarse_ln[k], ect, rq.xgam) * nf * 0.5 * (e0 + e1);
    }
    s_pe = block_sum(pe, s_red);
  }
  if (tid == 0) {
    b.rrcorr[rq.rr_corr_off + 2 * z] = s_ef / s_pe;
    b.rrcorr[rq.rr_corr_off + 2 * z + 1] = s_pf / s_pz / pow(1.5, rq.xgam);
  }
}

// relxill models with the nthcomp primary (relxillCp, relxilllpCp): one
// Kompaneets solve per (request, j) of work, j = 0 for the source's kTe (the
// nthcomp models keep kTe in the source frame, src/Relxill.cpp:197-200),
// j = 1 + z for the kTe lamp-post zone z sees (kTe times the zone's source ->
// disk energy shift, src/IonGradient.cpp:196-209, from the zone record as
// k_setup computes it).  Each solve writes its normalisation nf on the
// 500-bin grid and the energy flux of the normalised spectrum in 0.1-1000 keV
// (calc_xillver_normalization_change_source_to_disk, src/Xillspec.cpp:408-435;
// get_xillver_fluxcorr, :344-362); the source's also keeps its solution for
// k_primary.  First in a run: k_rrad_corr and k_setup read the results.
constexpr int NTH_THREADS_R = 256;

__global__ __launch_bounds__(NTH_THREADS_R) void k_nth(BatchDev b, const XillWork *__restrict__ work) {
  __shared__ NthLds L;
  __shared__ double s_red[NTH_THREADS_R];
  __shared__ double s_kte;
  const XillWork wk = work[blockIdx.x];
  const int req = wk.req, j = wk.zone, tid = threadIdx.x;
  const ReqParams &rq = b.req[req];
  if (rq.pre_status != 0) return;  // uniform over the block
  if (j == 0) {
    if (tid == 0) s_kte = rq.xect_src;
  } else {
    // the lamp-post rows in the solve's arrays (the solve overwrites them)
    lp_rows(b.lp, rq, L.a, L.b, L.c, L.d, tid, NTH_THREADS_R);
    if (tid == 0) {
      double ect;
      (void)zone_record(b.xt[rq.prim], rq, j - 1, true, L.a, L.b, L.c, L.d, b.lp.n_r, b.nr_fine, &ect);
      s_kte = ect;
    }
  }
  __syncthreads();
  const double kte = s_kte;
  const int nth = nth_solve(L, kte, rq.xgam, 0.05, tid, NTH_THREADS_R);  // kT_bb 0.05 keV (src/relutility.c:625-632)
  // the spectrum on the 500-bin grid (spec_nthcomp at z = 0), per bin in a
  const double nfac = nthcomp_normfac(L.x, L.c, nth, 0.0);
  const double *ce = b.cg.coarse;
  double sp = 0.0;
  for (int k = tid; k < N_ENER_COARSE; k += NTH_THREADS_R) {
    const double e0 = ce[k], e1 = ce[k + 1];
    const double f = nthcomp_bin(e0, e1, nthcomp_edge(L.x, L.c, nth, e0, 0.0), nthcomp_edge(L.x, L.c, nth, e1, 0.0),
                                 nfac);
    L.a[k] = f;
    if (e0 >= 0.1 && e0 <= 1000.0) sp += f * 0.5 * (e0 + e1) * 1e20 * 1.602177e-09;
  }
  const double nf = 1. / (block_sum(sp, s_red) / NORM_XILLVER_TABLE);
  double pe = 0.0;
  for (int k = tid; k < N_ENER_COARSE; k += NTH_THREADS_R) {
    const double e0 = ce[k], e1 = ce[k + 1];
    if (e0 >= 0.1 && e1 <= 1000.0) pe += L.a[k] * nf * 0.5 * (e0 + e1);
  }
  const double s_pe = block_sum(pe, s_red);
  if (tid == 0) b.nth[rq.nth_off + j] = NthRes{nf, s_pe};
  if (j == 0) {
    double *xo = b.nth_x + (size_t)rq.nth_sol * (NTH_N + 1);
    double *so = b.nth_s + (size_t)rq.nth_sol * NTH_N;
    for (int i = tid; i <= nth; i += NTH_THREADS_R) xo[i] = L.x[i];
    for (int i = tid; i < NTH_N; i += NTH_THREADS_R) so[i] = L.c[i];
    if (tid == 0) b.nth_n[rq.nth_sol] = nth;
  }
}

// Clears the run's counters; launched before k_setup on the batch's stream.
__global__ __launch_bounds__(64) void k_reset(int *ctr) {
  for (int i = threadIdx.x; i < N_COUNTERS; i += 64) ctr[i] = 0;
}

// ---------------------------------------------------------------------------
// integrate: the hot loop (calc_relline_profile, src/Relprofile.cpp:857-900)
// ---------------------------------------------------------------------------
namespace {

// The transfer function (and cosines) of one fine radius, lerped from its two
// table-node rows into a lane-private LDS row, and the FP64 edge norms.
// One row of a fine radius (2 n_g values), lerped from its two node rows into
// LDS; 8 values per round, so that a round's 16 loads are in flight together.
template <class TV>
__device__ inline void lerp_row(double ifac, const double *lo, const double *hi, int nq, TV *dst) {
  for (int q0 = 0; q0 < nq; q0 += 8) {
    double a[8], c[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int q = min(q0 + k, nq - 1);
      a[k] = lo[q];
      c[k] = hi[q];
    }
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (q0 + k < nq) dst[q0 + k] = (TV)lerp1_(ifac, a[k], c[k]);
  }
}

// A request's (a, mu0)-interpolated node rows (interpol_a_mu0,
// src/Relprofile.cpp:39-80), read from the table's four corners: row ir,
// column c (g* column c >> 1, branch c & 1) of the transfer function or of
// the emission-angle cosines.  k_setup computes the same lerp2f for the rows
// of the angle distribution, so every consumer sees the same bits.
struct NodeSrc {
  const float *a1, *a2;  // branch 0 and 1 arrays
  size_t s00, s10, s01, s11;
  double fa, fm;
  int ng;
};

__device__ inline NodeSrc node_src(const DevRelTable &T, const ReqNorm &rn, bool cos) {
  NodeSrc s;
  s.a1 = cos ? T.cosne1 : T.trff1;
  s.a2 = cos ? T.cosne2 : T.trff2;
  const size_t nq = (size_t)T.n_r * T.n_g;
  s.s00 = ((size_t)rn.ia * T.n_mu0 + rn.im) * nq;
  s.s10 = ((size_t)(rn.ia + 1) * T.n_mu0 + rn.im) * nq;
  s.s01 = ((size_t)rn.ia * T.n_mu0 + rn.im + 1) * nq;
  s.s11 = ((size_t)(rn.ia + 1) * T.n_mu0 + rn.im + 1) * nq;
  s.fa = rn.fa;
  s.fm = rn.fm;
  s.ng = T.n_g;
  return s;
}

// (the branch arrays are taken into locals before the select: a select
// between two members' addresses would put the struct in scratch memory)
__device__ inline double node_value(const NodeSrc &s, int ir, int c) {
  const float *a1 = s.a1, *a2 = s.a2;
  const float *arr = (c & 1) ? a2 : a1;
  const size_t q = (size_t)ir * s.ng + (c >> 1);
  return lerp2f(s.fa, s.fm, arr[s.s00 + q], arr[s.s10 + q], arr[s.s01 + q], arr[s.s11 + q]);
}

// Node rows kb .. kb + tot / nq - 1 (nq = 2 n_g values each, contiguous) built
// into LDS by the whole wavefront: every lane loads the four corners of PER
// values per round (4 PER loads in flight), then interpolates them; at most
// STAGE_MAX values.
constexpr int STAGE_MAX = 640;
template <int PER>
__device__ inline void stage_table_rows(const NodeSrc &s, int kb, int tot, int nq, double *dst, int lane) {
  int row = kb + lane / nq, col = lane % nq;
  const float *a1 = s.a1, *a2 = s.a2;
  for (int k0 = 0; k0 < STAGE_MAX / 64 && 64 * k0 < tot; k0 += PER) {
    float c00[PER], c10[PER], c01[PER], c11[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
      if (lane + 64 * (k0 + k) < tot) {
        const float *arr = (col & 1) ? a2 : a1;
        const size_t q = (size_t)row * s.ng + (col >> 1);
        c00[k] = arr[s.s00 + q];
        c10[k] = arr[s.s10 + q];
        c01[k] = arr[s.s01 + q];
        c11[k] = arr[s.s11 + q];
      }
      col += 64;
      while (col >= nq) {
        col -= nq;
        row++;
      }
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int idx = lane + 64 * (k0 + k);
      if (idx < tot) dst[idx] = lerp2f(s.fa, s.fm, c00[k], c10[k], c01[k], c11[k]);
    }
  }
}

// k_fixup: one lane's FP64 row values, straight from the table, for the g*
// brackets that the FP64 integral of the bin [e0, e1] reads (its points lie
// between the bin's clamped ends, whose brackets bound theirs; one bracket of
// margin either side) and for the brackets of the two edge norms; the other
// columns of the row are not read.
__device__ inline void load_radius_table(const BatchDev &b, const ReqNorm &rn, const RadGeo &r, int K, int limb,
                                         double e0, double e1, double *myT, double *myC, RadiusView &v) {
  v.gmin = r.gmin;
  v.gmax = r.gmax;
  v.del_g = 1. / (r.gmax - r.gmin);  // set_str_relbf, src/Relprofile.cpp:733
  v.limb = limb;
  const int ng = b.rt.n_g;
  auto bracket = [&](double e) {
    const double g = fmin(fmax((e - v.gmin) * v.del_g, GFAC_H), 1.0 - GFAC_H);
    return g_bracket(g, ng, v.gs_step, v.inv_step);
  };
  const NodeSrc ts = node_src(b.rt, rn, false);
  const NodeSrc cs = node_src(b.rt, rn, true);
  auto fill = [&](int ia, int ib) {  // columns of brackets ia .. ib
#pragma unroll 4
    for (int c = 2 * ia; c <= 2 * ib + 3; c++) {
      myT[c] = lerp1_(r.ifac_r, node_value(ts, K + 1, c), node_value(ts, K, c));
      if (limb) myC[c] = lerp1_(r.ifac_r, node_value(cs, K + 1, c), node_value(cs, K, c));
    }
  };
  fill(max(0, bracket(e0) - 1), min(ng - 2, bracket(e1) + 1));
  const EdgePt pL = edge_point(false, v), pU = edge_point(true, v);
  fill(pL.ind, pL.ind);
  fill(pU.ind, pU.ind);
  v.normL = edge_norm64(false, v);
  v.normU = edge_norm64(true, v);
}

__device__ inline const double *profile_grid(const BatchDev &b, const ReqParams *rq) {
  const int64_t off = rq->pgrid_off;
  return (off >= 0) ? b.pgrid + off : b.cg.ener;
}

}  // namespace

// Byte offset of k_integrate's coefficient slots and bin buffer (FP64 path)
// behind the lane rows (rows: 64 or 128 of 2 n_g + 1 values).
__host__ __device__ inline size_t integrate_rows_bytes(int ng, bool cos_rows, size_t elem) {
  const size_t b = (size_t)(cos_rows ? 2 : 1) * 64 * (2 * ng + 1) * elem;
  return (b + 15) & ~(size_t)15;
}

// Persistent wavefronts pull work items (request, chunk of 64 fine radii, bin
// segment) from the per-XCD ranges of the work list until every range is
// exhausted.  Lane l owns
// fine radius 64 rc + l: its transfer-function row sits in LDS and its window,
// weight and edge norms in registers.  For every bin of the segment each lane
// integrates the bin at its radius (zero outside its window) and parks the
// value in LDS; every SUM_BINS bins the wavefront turns round: 8 lanes per bin
// each sum 8 radii in FP64 (fixed order), a 3-step DPP tree adds the 8 partial
// sums and one lane per bin adds the result to the request's profile with one
// FP64 atomic.
//   MODE 0 (FP64): the parked values are the FP64 bin integrals; the
//   turn-round multiplies them by the radii's weights (64 slots in LDS).
//   MODE 1 (FP32 integrand, FP64 decisions and edges): each lane parks its
//   integral times its weight times the request's power of two
//   (ReqNorm.cmul) as FP32, and the atomic adds the sum times ReqNorm.cinv.
//   Bins whose Romberg stop test fell in the guard band are queued for the
//   FP64 re-check (k_fixup).
//   Item set-up: the node rows that the chunk's radii lerp between (nodes
//   kmin .. kmax + 1) are built in LDS from the table's four (a, mu0)
//   corners by the whole wavefront, stage_rows at a time, in one round of
//   loads each; each lane then lerps its row and the FP64 values of its edge
//   norms from LDS.  The staging area is the parked-value area, which is free
//   between items.
//   COS: the batch holds emission-angle rows (a request with limb
//   darkening); COS == 0 compiles the limb factors out.
template <int MODE, int COS>
__global__ __launch_bounds__(64, MODE == 0 ? 1 : 3) void k_integrate(BatchDev b, int stage_rows) {
  constexpr int cos_rows = COS;
  using TV = typename std::conditional<MODE == 0, double, float>::type;
  extern __shared__ double s_dyn[];
  const int lane = threadIdx.x;
  const int ng = b.rt.n_g;
  const int RS = 2 * ng + 1;  // odd row stride: lanes' rows start in distinct banks
  TV *rows = reinterpret_cast<TV *>(s_dyn);
  TV *myT = rows + lane * RS;
  TV *myC = cos_rows ? rows + (64 + lane) * RS : myT;
  // behind the rows: MODE 0 the 64 weights then the parked doubles, MODE 1 the parked floats
  double *s_coef = reinterpret_cast<double *>(reinterpret_cast<char *>(s_dyn) +
                                              integrate_rows_bytes(ng, cos_rows != 0, sizeof(TV)));
  TV *s_bins = (MODE == 0) ? reinterpret_cast<TV *>(s_coef + 64) : reinterpret_cast<TV *>(s_coef);
  // long items at [0, n_long), short ones at the back of the list
  const int n_long = min(b.ctr[CTR_ITEMS], b.item_cap);
  const int n_short = min(b.ctr[CTR_ITEMS_S], b.item_cap - n_long);
  const int nr = b.nr_fine;
  // turn-round roles: SUM_LANES lanes per bin; lane (col, part) sums radii
  // SUM_BINS part .. SUM_BINS part + SUM_BINS - 1 of bin col
  constexpr int SUM_LANES = 64 / SUM_BINS;
  const int col = lane / SUM_LANES, part = lane % SUM_LANES;
  RadiusView v;
  v.ng = ng;
  v.gs_step = (1.0 - 2 * GFAC_H) / (ng - 1);
  v.inv_step = 1.0 / v.gs_step;
  v.guard = b.guard;
  v.T64 = nullptr;
  v.C64 = nullptr;
  v.T32 = nullptr;
  v.C32 = nullptr;
  v.limb = 0;
  if (MODE == 0) {
    v.T64 = reinterpret_cast<const double *>(myT);
    v.C64 = reinterpret_cast<const double *>(myC);
  } else {
    v.T32 = reinterpret_cast<const float *>(myT);
    v.C32 = reinterpret_cast<const float *>(myC);
  }
  int xq = blockIdx.x % N_XQ, exhausted = 0;
  for (;;) {
    // next item of range xq (its share of the long items, then its share of
    // the short ones); an empty range sends the wavefront to the next one,
    // and once all N_XQ have been found empty it is done
    int k = -1;
    while (exhausted < N_XQ) {
      const int lo = (int)((int64_t)n_long * xq / N_XQ), hi = (int)((int64_t)n_long * (xq + 1) / N_XQ);
      const int los = (int)((int64_t)n_short * xq / N_XQ), his = (int)((int64_t)n_short * (xq + 1) / N_XQ);
      int kk = 0;
      if (lane == 0) kk = atomicAdd(&b.ctr[CTR_XQ + xq * XQ_STRIDE], 1);
      kk = __builtin_amdgcn_readfirstlane(kk);
      if (kk < hi - lo) {
        k = lo + kk;
        break;
      }
      if (kk < (hi - lo) + (his - los)) {
        k = b.item_cap - 1 - (los + kk - (hi - lo));
        break;
      }
      exhausted++;
      xq = (xq + 1) % N_XQ;
    }
    if (k < 0) break;
    const WorkItem it = b.items[k];
    const ReqParams *rq = b.req + it.req;
    const ReqNorm rn = b.rnorm[it.req];
    const int limb = COS ? rq->limb : 0;
    const int i = (it.rc & 0xffff) * RAD_CHUNK + lane;
    const int zone = it.rc >> 16;
    // the lane's radius: the stored record, and its geometry rebuilt from the
    // request's node arrays
    int rK = 0, ielo = 1, iehi = 0;
    double coef = 0.0;
    if (i < nr) {
      const RadRec pr = b.rad[(size_t)it.req * nr + i];
      if (rad_zone(pr) == zone) {  // (else a radius of another zone of the chunk)
        ielo = rad_ielo(pr);
        iehi = rad_iehi(pr);
      }
      rK = rad_K(pr);
      coef = pr.coef;
    }
    const bool has_win = ielo <= iehi;
    RadGeo r = {0.0, 0.0, 0.0};
    if (has_win) {
      const double *gn = b.rnode + (size_t)it.req * RAD_NODE_ARRAYS * b.rt.n_r;
      r = rad_geo(gn, gn + b.rt.n_r, gn + 2 * b.rt.n_r, fine_radius_at(rn.r1, rn.r2, i, nr), rK);
    }
    const int kmin = wave_min_i(has_win ? rK : INT_MAX);
    const int kmax = wave_max_i(has_win ? rK : -1);
    if (kmax >= 0) {
      // the lanes' rows from node rows kb .. ke + 1 built in LDS, stage_rows
      // at a time (lanes with K in [kb, ke] are served per round), for the
      // transfer function and then (limb) the cosines; the FP64 values at the
      // edge brackets are kept for the edge norms
      const int nq = 2 * ng;
      double *stage = reinterpret_cast<double *>(s_bins);
      const NodeSrc ts = node_src(b.rt, rn, false);
      const NodeSrc cs = node_src(b.rt, rn, true);
      EdgePt pL{}, pU{};
      if (has_win) {
        v.gmin = r.gmin;
        v.gmax = r.gmax;
        v.del_g = 1. / (r.gmax - r.gmin);  // set_str_relbf, src/Relprofile.cpp:733
        v.limb = limb;
        pL = edge_point(false, v);
        pU = edge_point(true, v);
      }
      for (int kb = kmin; kb <= kmax; kb += stage_rows - 1) {
        double tL[4] = {0, 0, 0, 0}, tU[4] = {0, 0, 0, 0}, cL[4] = {0, 0, 0, 0}, cU[4] = {0, 0, 0, 0};
        const int ke = min(kmax, kb + stage_rows - 2);
        const bool mine = has_win && rK >= kb && rK <= ke;
        const int tot = (ke - kb + 2) * nq;
        const int klo = mine ? rK + 1 - kb : 1, khi = mine ? rK - kb : 0;
        const double *lo = stage + (size_t)klo * nq;
        const double *hi = stage + (size_t)khi * nq;
        stage_table_rows<5>(ts, kb, tot, nq, stage, lane);
        __syncthreads();  // one wavefront: orders the staged rows before the lanes read them
        if (mine) {
          lerp_row<TV>(r.ifac_r, lo, hi, nq, myT);
#pragma unroll
          for (int k = 0; k < 4; k++) {
            tL[k] = lerp1_(r.ifac_r, lo[pL.ind * 2 + k], hi[pL.ind * 2 + k]);
            tU[k] = lerp1_(r.ifac_r, lo[pU.ind * 2 + k], hi[pU.ind * 2 + k]);
          }
        }
        __syncthreads();  // the next rows (or the bins) overwrite the staging area
        if (COS && limb) {
          stage_table_rows<5>(cs, kb, tot, nq, stage, lane);
          __syncthreads();
          if (mine) {
            lerp_row<TV>(r.ifac_r, lo, hi, nq, myC);
#pragma unroll
            for (int k = 0; k < 4; k++) {
              cL[k] = lerp1_(r.ifac_r, lo[pL.ind * 2 + k], hi[pL.ind * 2 + k]);
              cU[k] = lerp1_(r.ifac_r, lo[pU.ind * 2 + k], hi[pU.ind * 2 + k]);
            }
          }
          __syncthreads();
        }
        if (mine) {
          v.normL = edge_norm_at(pL, v, tL, cL);
          v.normU = edge_norm_at(pU, v, tU, cU);
          v.eL = (float)(2 * v.normL * (v.gmax - v.gmin));
          v.eU = (float)(2 * v.normU * (v.gmax - v.gmin));
        }
      }
    }
    const double w64 = has_win ? coef * rn.fac : 0.0;   // MODE 0 weight
    const double w32 = has_win ? coef * rn.cmul : 0.0;  // MODE 1 scaled weight
    const float w32f = (float)w32;
    if constexpr (MODE == 0) s_coef[lane] = w64;
    const double *E = profile_grid(b, rq);
    double *acc = profile_acc(b, *rq) + (size_t)zone * rq->n_e;
    double ehi = E[it.j0];
    double enext = E[min(it.j0 + 1, it.j1)];
    double ghi = (ehi - v.gmin) * v.del_g;  // unclamped g* of the bin's upper edge, carried
    EdgePair<TV> cy;  // integrand pair at the next bin's lower edge, when a Romberg bin left it
    cy.ok = false;
    for (int jg = it.j0; jg < it.j1; jg += SUM_BINS) {
      const int jn = min(SUM_BINS, it.j1 - jg);
      for (int jj = 0; jj < jn; jj++) {
        const int j = jg + jj;
        const double elo = ehi, glo = ghi;
        ehi = enext;
        enext = E[min(j + 2, it.j1)];  // next bin's upper edge, loaded a bin ahead
        ghi = (ehi - v.gmin) * v.del_g;
        TV val = 0;
        if (j >= ielo && j <= iehi) {
          if constexpr (MODE == 0) {
            val = integ_bin_g<0>(elo, ehi, glo, ghi, v, nullptr, &cy);
          } else {
            bool need64 = false;
            val = integ_bin32(elo, ehi, glo, ghi, v, &need64, &cy) * w32f;
            if (need64) {
              const int slot = atomicAdd(&b.ctr[CTR_FIX], 1);
              if (slot < b.fix_cap) {
                FixEntry fe;
                fe.req = it.req;
                fe.rad = i;
                fe.bin = j;
                fe.pad = 0;
                fe.v32 = (double)val * rn.cinv;  // this radius' share of the bin's atomic
                b.fix[slot] = fe;
              } else {
                atomicOr(&b.ctr[CTR_FIX_OVF], 1);
              }
            }
          }
        } else {
          cy.ok = false;
        }
        s_bins[jj * SUM_STRIDE + lane] = val;
      }
      __syncthreads();  // one wavefront: orders the parked values before the turn-round
      // FP64 sum over the 64 radii of bin jg + col: SUM_BINS radii per lane,
      // then the SUM_LANES lanes of the bin in a fixed DPP tree
      const TV *src = s_bins + col * SUM_STRIDE + part * SUM_BINS;
      double sum = 0.0;
      if constexpr (MODE == 0) {
        const double *cf = s_coef + part * SUM_BINS;
#pragma unroll
        for (int q = 0; q < SUM_BINS; q++) sum = fma(src[q], cf[q], sum);
      } else {
#pragma unroll
        for (int q = 0; q < SUM_BINS; q++) sum += (double)src[q];
      }
      sum += dpp_d<0xB1>(sum);   // quad_perm [1,0,3,2]
      sum += dpp_d<0x4E>(sum);   // quad_perm [2,3,0,1]
      sum += dpp_d<0x141>(sum);  // row_half_mirror: every lane of a half row holds its sum
      if constexpr (SUM_LANES == 16) sum += dpp_d<0x140>(sum);  // row_mirror: the whole row
      if (MODE == 1) sum *= rn.cinv;
      if (part == 0 && col < jn) unsafeAtomicAdd(&acc[jg + col], sum);
      __syncthreads();  // the next bins overwrite the parked values
    }
  }
}

// FP64 re-check of guarded FP32 bins, one entry per lane.
__global__ __launch_bounds__(64) void k_fixup(BatchDev b, int cos_rows) {
  extern __shared__ double s_dyn[];
  const int lane = threadIdx.x;
  const int ng = b.rt.n_g;
  const int RS = 2 * ng + 1;
  double *myT = s_dyn + lane * RS;
  double *myC = cos_rows ? s_dyn + (64 + lane) * RS : myT;
  const int n = min(b.ctr[CTR_FIX], b.fix_cap);
  RadiusView v;
  v.ng = ng;
  v.gs_step = (1.0 - 2 * GFAC_H) / (ng - 1);
  v.inv_step = 1.0 / v.gs_step;
  v.guard = b.guard;
  v.T64 = myT;
  v.C64 = myC;
  v.T32 = nullptr;
  v.C32 = nullptr;
  for (int e = blockIdx.x * 64 + lane; e < n; e += gridDim.x * 64) {
    const FixEntry fe = b.fix[e];
    const ReqParams *rq = b.req + fe.req;
    const RadRec pr = b.rad[(size_t)fe.req * b.nr_fine + fe.rad];
    const double *gn = b.rnode + (size_t)fe.req * RAD_NODE_ARRAYS * b.rt.n_r;
    const ReqNorm &rn = b.rnorm[fe.req];
    const RadGeo r = rad_geo(gn, gn + b.rt.n_r, gn + 2 * b.rt.n_r, fine_radius_at(rn.r1, rn.r2, fe.rad, b.nr_fine),
                             rad_K(pr));
    const double *E = profile_grid(b, rq);
    load_radius_table(b, b.rnorm[fe.req], r, rad_K(pr), rq->limb, E[fe.bin], E[fe.bin + 1], myT, myC, v);
    const double val = integ_bin<0>(E[fe.bin], E[fe.bin + 1], v) * (pr.coef * b.rnorm[fe.req].fac);
    unsafeAtomicAdd(&profile_acc(b, *rq)[(size_t)rad_zone(pr) * rq->n_e + fe.bin], val - fe.v32);
  }
}

// ---------------------------------------------------------------------------
// finalize: renormalisation (renorm_relline_profile, src/Relprofile.cpp:749-796)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(SETUP_THREADS) void k_finalize(BatchDev b) {
  __shared__ double s_red[SETUP_THREADS];
  const int req = blockIdx.x;
  const int tid = threadIdx.x;
  const ReqParams rq = b.req[req];
  if (rq.model == MOD_XILLVER) return;  // no profile; k_primary writes the output
  const int n_e = rq.n_e;
  double *dst = profile_acc(b, rq);
  const int st = b.status[req];
  if (st != 0) {
    for (int j = tid; j < n_e * rq.n_zones; j += SETUP_THREADS) dst[j] = 0.0;
    return;
  }
  if (rq.prof_off < 0) {  // line models; convolution profiles are finished by k_conv
    const double *E = b.pgrid + rq.pgrid_off;
    const BinRange rc = b.reach[req];
    double loc = 0.0;
    for (int j = tid; j < n_e; j += SETUP_THREADS) {
      double v = 0.0;
      if (j >= rc.lo && j <= rc.hi) v = dst[j] / (0.5 * (E[j] + E[j + 1]));
      dst[j] = v;
      loc += v;
    }
    const double sum = block_sum(loc, s_red);
    if (rq.do_renorm) {
      const double f = rq.norm / sum;
      for (int j = tid; j < n_e; j += SETUP_THREADS) dst[j] *= f;
    }
  }
  for (int z = 0; rq.n_incl > 0 && z < rq.n_zones; z++) {  // each zone's distribution sums to 1
    double *d = b.dist + rq.dist_off + (size_t)z * rq.n_incl;
    double ds = 0.0;
    for (int k = tid; k < rq.n_incl; k += SETUP_THREADS) ds += d[k];
    const double dsum = block_sum(ds, s_red);
    for (int k = tid; k < rq.n_incl; k += SETUP_THREADS) d[k] /= dsum;
  }
}

// ---------------------------------------------------------------------------
// convolution
// ---------------------------------------------------------------------------
namespace {

// LDS index of complex element i: one pad slot every 16 elements, so that the
// stride-16 writes of a radix-16 pass fall into distinct banks.
__device__ inline int P(int i) { return i + (i >> 4); }

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * (S i), S = -1 (forward) or +1 (inverse)
template <int S>
__device__ inline double2 mul_si(double2 a) {
  return S > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// exp(S 2 pi i M / 16), M known at compile time
template <int S, int M>
__device__ inline double2 w16() {
  constexpr double c[5] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173, 0.0};
  constexpr int q = (M & 15) >> 2, r = M & 3;
  constexpr double cs = q == 0 ? c[r] : q == 1 ? -c[4 - r] : q == 2 ? -c[r] : c[4 - r];
  constexpr double sn = q == 0 ? c[4 - r] : q == 1 ? c[r] : q == 2 ? -c[4 - r] : -c[r];
  return make_double2(cs, S * sn);
}

template <int S>
__device__ __forceinline__ void dft4(double2 &a0, double2 &a1, double2 &a2, double2 &a3) {
  const double2 s02 = cadd(a0, a2), d02 = csub(a0, a2);
  const double2 s13 = cadd(a1, a3), d13 = mul_si<S>(csub(a1, a3));
  a0 = cadd(s02, s13);
  a2 = csub(s02, s13);
  a1 = cadd(d02, d13);
  a3 = csub(d02, d13);
}

template <int S, int N2, int K1>
__device__ inline void tw16(double2 &x) {
  if (N2 * K1 != 0) x = cmul(x, w16<S, N2 * K1>());
}

// In-register DFT of length R (2, 4, 8 or 16), natural order in and out.
template <int R, int S>
__device__ __forceinline__ void dft(double2 *v) {
  if (R == 2) {
    const double2 t = v[0];
    v[0] = cadd(t, v[1]);
    v[1] = csub(t, v[1]);
  } else if (R == 4) {
    dft4<S>(v[0], v[1], v[2], v[3]);
  } else if (R == 8) {
    // n = 2 n1 + n2, k = k1 + 4 k2: two DFT4 (even, odd), twiddle, butterflies
    dft4<S>(v[0], v[2], v[4], v[6]);
    dft4<S>(v[1], v[3], v[5], v[7]);
    tw16<S, 1, 2>(v[3]);
    tw16<S, 1, 4>(v[5]);
    tw16<S, 1, 6>(v[7]);
    const double2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    const double2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
  } else {
    // n = 4 n1 + n2, k = k1 + 4 k2: DFT4 over n1 for each n2 (result k1 in
    // slot n2 + 4 k1), twiddle w16^(n2 k1), DFT4 over n2 for each k1 (result
    // k2 in slot n2 = k2 of row k1), then the transpose to natural order.
    dft4<S>(v[0], v[4], v[8], v[12]);
    dft4<S>(v[1], v[5], v[9], v[13]);
    dft4<S>(v[2], v[6], v[10], v[14]);
    dft4<S>(v[3], v[7], v[11], v[15]);
    tw16<S, 1, 1>(v[5]);
    tw16<S, 1, 2>(v[9]);
    tw16<S, 1, 3>(v[13]);
    tw16<S, 2, 1>(v[6]);
    tw16<S, 2, 2>(v[10]);
    tw16<S, 2, 3>(v[14]);
    tw16<S, 3, 1>(v[7]);
    tw16<S, 3, 2>(v[11]);
    tw16<S, 3, 3>(v[15]);
    dft4<S>(v[0], v[1], v[2], v[3]);
    dft4<S>(v[4], v[5], v[6], v[7]);
    dft4<S>(v[8], v[9], v[10], v[11]);
    dft4<S>(v[12], v[13], v[14], v[15]);
    // slot 4 k1 + k2 holds X[k1 + 4 k2]
    double2 t[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++)
#pragma unroll
      for (int k2 = 0; k2 < 4; k2++) t[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = t[k];
  }
}

// One Stockham radix-R pass over N complex values in LDS (padded layout):
// thread j loads x[j + r N/R] r = 0..R-1, twiddles by w^(r (j mod Ns)),
// w = exp(S 2 pi i / (Ns R)) (table twp[r Ns + k], forward sign, conjugated
// for S = +1), transforms, and stores at (j / Ns) Ns R + j mod Ns + r Ns.
// PER j values per thread; reads finish before any write.
template <int R, int S, int PER>
__device__ __forceinline__ void fft_pass(double2 *buf, int N, int Ns, const double2 *twp) {
  const int J = N / R;
  double2 v[PER][R];
  const int lgNs = __ffs(Ns) - 1;  // Ns is a power of two
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int j = threadIdx.x + u * blockDim.x;
    if (j < J) {
      const int k = j & (Ns - 1);
#pragma unroll
      for (int r = 0; r < R; r++) {
        double2 x = buf[P(j + r * J)];
        if (r > 0) {
          double2 w = twp[r * Ns + k];
          if (S > 0) w.y = -w.y;
          x = cmul(x, w);
        }
        v[u][r] = x;
      }
      dft<R, S>(v[u]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int j = threadIdx.x + u * blockDim.x;
    if (j < J) {
      const int k = j & (Ns - 1);
      const int base = ((j >> lgNs) << lgNs) * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) buf[P(base + r * Ns)] = v[u][r];
    }
  }
  __syncthreads();
}

// The block has at least N/8 threads, so a pass holds at most 8 values per
// thread: PER = 1 (R = 8), <= 2 (R = 4), <= 4 (R = 2); radix 16 keeps PER = 1.
template <int R, int S>
__device__ __forceinline__ void fft_pass_any(double2 *buf, int N, int Ns, const double2 *twp) {
  const int per = (N / R + blockDim.x - 1) / blockDim.x;
  if (per <= 1) fft_pass<R, S, 1>(buf, N, Ns, twp);
  else if (R <= 8 && per <= 2) fft_pass<R, S, (R <= 8 ? 2 : 1)>(buf, N, Ns, twp);
  else if (R <= 4 && per <= 4) fft_pass<R, S, (R <= 4 ? 4 : 1)>(buf, N, Ns, twp);
  else fft_pass<R, S, (R <= 2 ? 8 : 1)>(buf, N, Ns, twp);
}

// Unnormalised complex FFT of N (power of two, 2..8192) values in LDS,
// natural order in and out; radix-8 passes, then one of radix 4 or 2 (the
// block has at least N/8 threads).  twp holds the passes' tables one after
// the other (R Ns entries each), in the same plan as the host builds them.
template <int S>
__device__ __forceinline__ void lds_fft(double2 *buf, int N, const double2 *twp) {
  int Ns = 1;
  while (Ns < N) {
    const int left = N / Ns;
    if (left % 8 == 0) {
      fft_pass_any<8, S>(buf, N, Ns, twp);
      twp += 8 * Ns;
      Ns *= 8;
    } else if (left % 4 == 0) {
      fft_pass_any<4, S>(buf, N, Ns, twp);
      twp += 4 * Ns;
      Ns *= 4;
    } else {
      fft_pass_any<2, S>(buf, N, Ns, twp);
      twp += 2 * Ns;
      Ns *= 2;
    }
  }
}

// Block sum in a fixed order: FP64 wave trees, then the waves in order.
__device__ inline double block_sum4(double v, double *s_w) {
  const double w = wave_sum(v);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) s_w[wave] = w;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += s_w[k];
  __syncthreads();
  return t;
}

// Two block sums at once (one pair of barriers).
__device__ inline void block_sum2(double a, double c, double *out, double (*s_w)[CONV_THREADS_MAX / 64]) {
  const double wa = wave_sum(a), wc = wave_sum(c);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_w[0][wave] = wa;
    s_w[1][wave] = wc;
  }
  __syncthreads();
  double ta = 0.0, tc = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
    ta += s_w[0][k];
    tc += s_w[1][k];
  }
  __syncthreads();
  out[0] = ta;
  out[1] = tc;
}

// calcFFTNormFactor band (src/Relbase.cpp:205): bins with E_lo >= 0.01 and
// E_hi < 1000, a contiguous index range of the convolution grid
__device__ inline bool in_band(const DevConvGrid &cg, int i) { return i >= cg.band_lo && i < cg.band_hi; }



}  // namespace

// Convolution of one request (relconv_kernel minus relbase, src/Relbase.cpp:257-289).
// One complex FFT carries both real inputs (input spectrum in the real part,
// rotated profile in the imaginary part); their spectra are separated, multiplied
// and inverted with a half-length complex FFT.  Grid-walking phases give each
// wavefront a contiguous run of bins, one bin per lane, so that neighbouring
// lanes read neighbouring addresses.
__global__ __launch_bounds__(CONV_THREADS_MAX) void k_conv(BatchDev b, const XillWork *__restrict__ work) {
  extern __shared__ double2 s_buf[];  // P(n) complex
  __shared__ double s_w[2][CONV_THREADS_MAX / 64];
  const int NT = blockDim.x;
  const int req = work[blockIdx.x].req;
  const int zone = work[blockIdx.x].zone;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, nwave = NT >> 6;
  const ReqParams rq = b.req[req];
  const DevConvGrid cg = b.cg;
  const int n = cg.n;
  const int m = n >> 1;
  double *out = b.out + rq.flux_off;
  // output grid: the caller's (for relxill already redshifted on the host)
  const double *eo = b.ener_in + rq.ener_off;
  const int n_o = rq.n_in;
  if (b.status[req] != 0) {
    for (int j = tid; j < n_o; j += NT) out[j] = 0.0;
    return;
  }
  const double *Ec = cg.ener;
  const double *cf = cg.cfac;
  const double2 *tw = reinterpret_cast<const double2 *>(cg.tw);
  const bool angular = (rq.model == MOD_RELXILL_CONV);
  const bool relxill = (rq.model == MOD_RELXILL || rq.model == MOD_RELXILLLP);
  // input: the caller's spectrum, or for relxill each zone's angle-weighted
  // xillver spectrum on the table grid (k_xill)
  const double *ein = relxill ? b.xt[rq.prim].ener : eo;
  const int n_in = relxill ? b.xt[rq.prim].n_ener : n_o;
  const DevReflSet *rs = angular ? &b.refl_sets[rq.refl_set] : nullptr;
  auto ge_in = [&](int k) { return ein[k]; };
  const BinRange rc = b.reach[req];
  // relxill: the reflection's scale (add_primary_spectrum,
  // src/PrimarySource.cpp:66-125): |refl_frac| for the broken power law; for
  // the lamp post |refl_frac| (times the lamp post's own reflection fraction
  // when refl_frac is a boost) over that fraction.  The primary is added by
  // k_primary.
  double refl_scale = 1.0;
  if (relxill) {
    if (rq.emis_type != EMIS_LP) {
      refl_scale = fabs(rq.refl_frac);
    } else {
      const LpRefl lr = b.refl[req];
      double rfi = rq.refl_frac;
      if (rq.xboost) rfi *= lr.refl_frac;
      refl_scale = (fabs(rfi)) / lr.refl_frac;
    }
  }
  auto o_at = [&](int i) -> double & {
    double2 &z = s_buf[P(i >> 1)];
    return (i & 1) ? z.y : z.x;
  };

  // this block's zone (relxill requests have one block per zone)
  const double *prof = b.prof + rq.prof_off + (size_t)zone * n;
  const double *fin = relxill ? b.xspec + rq.xspec_off + (size_t)zone * n_in : b.flux_in + rq.flux_off;
  const double *dist = b.dist + rq.dist_off + (size_t)zone * rq.n_incl;
  // 1. x = rebinned input * E/dE (real part), y = rotated profile * E/dE (imag part)
  double sx = 0.0;
  const int nb = n / nwave;  // bins per wavefront (a multiple of 64)
  if (!angular) {
    // _rebin_spectrum of the input, 4 conv bins per lane at a time: the edges
    // around the guessed brackets of all 4 are loaded together, then their
    // end fluxes; the exact search runs only where a guess misses
    const GridGuess gi = grid_guess(ein, n_in);
    for (int q0 = 0; q0 < nb; q0 += 4 * 64) {
      double elo[4], ehi[4], al[4], bl[4], ah[4], bh[4];
      int kl[4], kh[4], i4[4];
      bool use[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const bool valid = q0 + 64 * u < nb;
        i4[u] = wave * nb + (valid ? q0 + 64 * u : 0) + lane;
        elo[u] = Ec[i4[u]];
        ehi[u] = Ec[i4[u] + 1];
        // overlap with the input grid (the first test of _rebin_spectrum)
        use[u] = valid && gi.first <= ehi[u] && gi.last >= elo[u];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        kl[u] = guess_index(elo[u], gi, n_in);
        kh[u] = guess_index(ehi[u], gi, n_in);
        if (use[u]) {
          al[u] = ein[kl[u]];
          bl[u] = ein[kl[u] + 1];
          ah[u] = ein[kh[u]];
          bh[u] = ein[kh[u] + 1];
        }
      }
      int imin[4], imax[4];
      double fa[4], fb[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (!use[u]) continue;
        const int clo = count_le_at(ge_in, n_in, elo[u], gi, kl[u], al[u], bl[u]);
        const int chi = count_le_at(ge_in, n_in, ehi[u], gi, kh[u], ah[u], bh[u]);
        imin[u] = clo > 0 ? clo - 1 : 0;
        imax[u] = chi < n_in ? chi : n_in;
        if (imax[u] > 0) imax[u]--;
        if (imin[u] >= n_in) {
          use[u] = false;
          continue;
        }
        if (imin[u] != kl[u]) {  // a missed guess or a clipped end: the edges of imin
          al[u] = ein[imin[u]];
          bl[u] = ein[imin[u] + 1];
        }
        if (imax[u] != kh[u]) {
          ah[u] = ein[imax[u]];
          bh[u] = ein[imax[u] + 1];
        }
        fa[u] = fin[imin[u]];
        fb[u] = fin[imax[u]];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (q0 + 64 * u >= nb) continue;
        double xr = 0.0;
        if (use[u]) {  // rebin_bin with the edges and end fluxes in registers
          const double el = (elo[u] < al[u]) ? al[u] : elo[u];
          const double eh = (ehi[u] > bh[u]) ? bh[u] : ehi[u];
          const double ra = rcp64(bl[u] - al[u]);
          if (imax[u] == imin[u]) {
            xr = (eh - el) * ra * fa[u];
          } else {
            const double dmin = (bl[u] - el) * ra;
            const double dmax = (eh - ah[u]) * rcp64(bh[u] - ah[u]);
            xr += fa[u] * dmin + fb[u] * dmax;
            for (int jj = imin[u] + 1; jj <= imax[u] - 1; jj++) xr += fin[jj];
          }
        }
        const int i = i4[u];
        if (elo[u] >= 0.01 && ehi[u] < 1000.0) sx += xr;  // calcFFTNormFactor band (src/Relbase.cpp:205)
        s_buf[P(i)].x = xr * cf[i];
      }
    }
  } else {
    for (int i = tid; i < n; i += NT) s_buf[P(i)].x = 0.0;
  }
  // the profile as photons per bin, renormalised (renorm_relline_profile,
  // src/Relprofile.cpp:757-779); bins no window reaches are zero
  const int per = n / NT;  // <= 8: the block has at least n/8 threads
  double ph[8];
  double sph = 0.0;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int i = tid + q * NT;
    ph[q] = (q < per && i >= rc.lo && i <= rc.hi) ? prof[i] * cg.inv_emid[i] : 0.0;
    sph += ph[q];
  }
  double s2[2];
  block_sum2(sx, sph, s2, s_w);
  double sum_xill = s2[0];
  const double f = rq.do_renorm ? rq.norm / s2[1] : 1.0;
  double sr = 0.0, sa = 0.0;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int i = tid + q * NT;
    if (q < per) {
      const double pv = ph[q] * f;
      if (in_band(cg, i)) sr += pv;
      sa += pv;
      const int irot = (i - cg.i1kev + n) & (n - 1);
      s_buf[P(irot)].y = pv * cf[i];
    }
  }
  // (barrier inside: the buffer is complete)
  double sum_rel, sum_prof = 1.0;
  if (angular || relxill) {
    double s3[2];
    block_sum2(sr, sa, s3, s_w);
    sum_rel = s3[0];
    sum_prof = s3[1];
  } else {
    sum_rel = block_sum4(sr, s_w[0]);
  }
  if (angular) {
    sum_xill = 0.0;
    for (int k = 0; k < rs->n_incl; k++) sum_xill += dist[k] * rs->bandsum[k];
  }
  // a zone whose profile sums below 1e-12 is left out (src/Relxill.cpp:453-456)
  const bool zone_used = !(sum_prof < 1e-12);
  if (relxill && !zone_used) return;  // (block-uniform)
  // 2. forward complex FFT of x + i y
  lds_fft<-1>(s_buf, n, reinterpret_cast<const double2 *>(cg.twp_fwd));
  // 3. X = (W_k + conj W_{n-k})/2, Y = (W_k - conj W_{n-k})/(2i), P = X Y (k = 0..m);
  //    then pack for the half-length inverse real FFT:
  //    Z_k = (P_k + conj P_{m-k}) + i (P_k - conj P_{m-k}) e^{+2 pi i k/n}.
  //    The pair {k, m-k} reads slots k, n-k, m-k, m+k and writes slots k and
  //    m-k, which no other pair touches: no staging across a barrier.
  for (int k = tid; k <= (m >> 1); k += NT) {
    double2 Pk[2];
    const int ks[2] = {k, m - k};
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int kk = ks[u];
      const double2 wk = s_buf[P(kk)];
      const double2 wn = s_buf[P((n - kk) & (n - 1))];
      double2 Y = make_double2(0.5 * (wk.y + wn.y), -0.5 * (wk.x - wn.x));
      double2 X;
      if (angular) {
        X = make_double2(0.0, 0.0);
        for (int a = 0; a < rs->n_incl; a++) {
          const double *F = rs->fft + ((size_t)a * (m + 1) + kk) * 2;
          X.x += dist[a] * F[0];
          X.y += dist[a] * F[1];
        }
      } else {
        X = make_double2(0.5 * (wk.x + wn.x), 0.5 * (wk.y - wn.y));
      }
      Pk[u] = cmul(X, Y);
    }
    if (k == 0) {  // c2r ignores the imaginary parts of P_0 and P_m
      Pk[0].y = 0.0;
      Pk[1].y = 0.0;
    }
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int kk = ks[u];
      if (kk >= m || (u == 1 && kk == k)) continue;  // Z_m is not needed; k = m/2 once
      const double2 p0 = Pk[u], p1 = Pk[1 - u];
      const double2 A = make_double2(p0.x + p1.x, p0.y - p1.y);
      const double2 Bd = make_double2(p0.x - p1.x, p0.y + p1.y);
      double2 w = tw[kk];
      w.y = -w.y;  // e^{+2 pi i kk/n}
      const double2 Bw = cmul(Bd, w);
      s_buf[P(kk)] = make_double2(A.x - Bw.y, A.y + Bw.x);
    }
  }
  __syncthreads();
  // 4. half-length inverse FFT: z_j = out_{2j} + i out_{2j+1}
  lds_fft<+1>(s_buf, m, reinterpret_cast<const double2 *>(cg.twp_inv));
  double sc = 0.0;
#pragma unroll 8
  for (int i = tid; i < n; i += NT) {
    const double v = o_at(i) * cg.cfinv[i];
    o_at(i) = v;
    if (in_band(cg, i)) sc += v;
  }
  const double sum_conv = block_sum4(sc, s_w[0]);
  // normalizeFFTOutput (src/Relbase.cpp:199-222): applied to the rebinned
  // values below (the rebin is linear), which saves a pass over the buffer
  double fnorm = sum_rel * sum_xill / sum_conv;
  if (angular && !zone_used) fnorm = 0.0;
  if (relxill) fnorm *= refl_scale;
  // the conv grid edges go to the upper part of the buffer, free after the
  // half-length FFT (its values occupy slots below P(m))
  double *s_e = reinterpret_cast<double *>(s_buf + P(m));
#pragma unroll 8
  for (int i = tid; i <= n; i += NT) s_e[i] = Ec[i];
  __syncthreads();
  // 5. rebin onto the caller grid; relconv zeroes outside 0.01-1000 keV;
  // relxill adds its zones up
  const GridGuess gc = grid_guess(Ec, n);
  auto ge_c = [&](int k) { return s_e[k]; };
  const int nbo = (n_o + nwave - 1) / nwave;
  const int jend = min(n_o, (wave + 1) * nbo);
  const bool relconv = !angular && !relxill;
#pragma unroll 2
  for (int q = 0; q < nbo; q += 64) {
    const int jb = wave * nbo + q + lane;
    if (jb < jend) {
      const double elo = eo[jb], ehi = eo[jb + 1];
      const int clo = count_le(ge_c, n, elo, gc);
      const int chi = count_le(ge_c, n, ehi, gc);
      double v = fnorm * rebin_bin(elo, ehi, clo, chi, ge_c, n, [&](int i) { return o_at(i); });
      if (relconv && (ehi < 0.01 || elo > 1000.0)) v = 0.0;
      if (relxill) unsafeAtomicAdd(&out[jb], v);  // the zones add up (k_setup cleared out)
      else out[jb] = v;
    }
  }
}

// relxill: the primary added to the reflection (add_primary_spectrum,
// src/PrimarySource.cpp:66-125) for refl_frac >= 0: the cut-off power law with
// the source's Ecut shifted to the observer (spec_cutoffpl,
// src/Xillspec.cpp:215-241) on the convolution grid, normalised as the
// xillver incident spectrum at the source (calcNormWrtXillverTableSpec on the
// 500-bin grid, src/PrimarySource.h:238-292), times the lamp post's flux boost,
// rebinned onto the caller grid.  One workgroup per request, after k_conv.
__global__ __launch_bounds__(PRIM_THREADS) void k_primary(BatchDev b, const int *__restrict__ reqs) {
  __shared__ double s_w[PRIM_THREADS / 64];
  const int req = reqs[blockIdx.x];
  const int tid = threadIdx.x, NT = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nwave = NT >> 6;
  const ReqParams &rq = b.req[req];
  const bool xill = rq.model == MOD_XILLVER;
  if (xill && b.status[req] != 0) {
    for (int j = tid; j < rq.n_in; j += NT) b.out[rq.flux_off + j] = 0.0;
    return;
  }
  if (!xill && (b.status[req] != 0 || !(rq.refl_frac >= 0))) return;
  const DevConvGrid &cg = b.cg;
  const int n = cg.n;
  const double ect = rq.xect_src, gam = rq.xgam;
  double prim_scale = 1.0;
  if (rq.emis_type == EMIS_LP) {  // flux_boost_source_to_observer, src/PrimarySource.h:84-93
    prim_scale = b.refl[req].f_inf_rest / 0.5 * pow(rq.g_so, gam);
    if (rq.beta > 1e-4) {
      const double dop = sqrt(1.0 - rq.beta * rq.beta) / (1.0 + rq.beta * cos(M_PI - rq.incl));
      prim_scale *= pow(dop, 2);
    }
  }
  // nthcomp: the source's solve (k_nth) observed at redshift 1/g - 1
  // (spec_nthcomp, src/Xillspec.cpp:243-253) with the norm of its spectrum
  // at the source
  const bool nth = rq.prim == PRIM_NTHCOMP;
  const int sol = nth ? rq.nth_sol : 0;
  const double *nx = b.nth_x + (size_t)sol * (NTH_N + 1);
  const double *ns = b.nth_s + (size_t)sol * NTH_N;
  const int nn = nth ? b.nth_n[sol] : 0;
  const double zo = 1 / rq.g_so - 1;
  const double nfac = nth ? nthcomp_normfac(nx, ns, nn, zo) : 0.0;
  // the blackbody (relxillNS) takes no energy shift (calc_primary_spectrum,
  // src/Xillspec.cpp:293-295)
  const bool bb = rq.prim == PRIM_BBODY;
  const double nf = nth  ? b.nth[rq.nth_off].nf
                    : bb ? 1. / (block_sum4(bbody_norm_part(cg, rq.xktbb, tid, NT), s_w) / NORM_XILLVER_TABLE)
                         : 1. / (block_sum4(xill_norm_part(cg, ect, gam, tid, NT), s_w) / NORM_XILLVER_TABLE);
  // the primary on the convolution grid, evaluated where the rebin reads it
  // (one or two grid bins per output bin on caller grids finer than the
  // convolution grid; every grid bin of the output bin, each with its
  // exponential, on coarser ones): no staging buffer
  const double *ce = cg.ener;
  const double *cl = cg.ln_emid;
  const double ect_o = ect * rq.g_so;
  auto prim_at = [&](int i) {
    const double v = nth  ? nthcomp_bin(ce[i], ce[i + 1], nthcomp_edge(nx, ns, nn, ce[i], zo),
                                        nthcomp_edge(nx, ns, nn, ce[i + 1], zo), nfac)
                     : bb ? bbody_bin(ce[i], ce[i + 1], rq.xktbb)
                          : cutoffpl_bin(ce[i], ce[i + 1], cl[i], ect_o, gam);
    return (v * nf) * prim_scale;
  };
  const GridGuess gc = grid_guess(ce, n);
  auto ge_c = [&](int k) { return ce[k]; };
  const double *eo = b.ener_in + rq.ener_off;
  double *out = b.out + rq.flux_off;
  const int n_o = rq.n_in;
  const int nbo = (n_o + nwave - 1) / nwave;
  const int jend = min(n_o, (wave + 1) * nbo);
  if (xill) {
    // the xillver models: the table spectrum times 0.5 cos(Incl)
    // (norm_xillver_spec, src/Xillspec.cpp:536-545) rebinned onto the
    // (redshifted) caller grid, times |refl_frac|, plus the normalised primary
    // evaluated on the caller grid itself for refl_frac >= 0
    // (add_primary_component, src/Relbase.cpp:294-344)
    const DevXillTable &xt = b.xt[rq.prim];
    const int nxt = xt.n_ener;
    const double *xs = b.xspec + rq.xspec_off;
    const double slab = rq.norm, rf = rq.refl_frac;
    const GridGuess gx = grid_guess(xt.ener, nxt);
    auto ge_x = [&](int k) { return xt.ener[k]; };
    auto refl_at = [&](int k) { return xs[k] * slab; };
    for (int q = 0; q < nbo; q += 64) {
      const int jb = wave * nbo + q + lane;
      if (jb < jend) {
        const double elo = eo[jb], ehi = eo[jb + 1];
        const int xlo = count_le(ge_x, nxt, elo, gx);
        const int xhi = count_le(ge_x, nxt, ehi, gx);
        double v = rebin_bin(elo, ehi, xlo, xhi, ge_x, nxt, refl_at) * fabs(rf);
        if (rf >= 0) {
          const double p = nth  ? nthcomp_bin(elo, ehi, nthcomp_edge(nx, ns, nn, elo, 0.0),
                                              nthcomp_edge(nx, ns, nn, ehi, 0.0), nfac)
                           : bb ? bbody_bin(elo, ehi, rq.xktbb)
                                : cutoffpl_bin(elo, ehi, log(0.5 * (elo + ehi)), ect, gam);
          v += p * nf;
        }
        out[jb] = v;
      }
    }
    return;
  }
  for (int q = 0; q < nbo; q += 64) {
    const int jb = wave * nbo + q + lane;
    if (jb < jend) {
      const double elo = eo[jb], ehi = eo[jb + 1];
      const int clo = count_le(ge_c, n, elo, gc);
      const int chi = count_le(ge_c, n, ehi, gc);
      out[jb] += rebin_bin(elo, ehi, clo, chi, ge_c, n, prim_at);
    }
  }
}

// relxill: the xillver spectrum of the request's parameters for every
// emission angle (interp_5d_tab_incl, src/xilltable.c:812-876: 16 corners of
// the four interpolated parameters, in the reference's order and with its
// weight products) weighted by the angle distribution (calc_xillver_angdep,
// src/Xillspec.cpp:557-573), for one-zone requests.  XILL_BINS bins per
// thread; every table read is coalesced across the block.  Angles with zero
// weight are skipped (they add exactly 0).  The zone's Ecut bracket, angle
// distribution and normalisation change come from its zone record.
__global__ __launch_bounds__(XILL_THREADS, 4) void k_xill1(BatchDev b, const XillWork *__restrict__ work, int chunks) {
  const XillWork wk = work[blockIdx.x / chunks];
  const int req = wk.req;
  if (b.status[req] != 0) return;
  const ReqParams &rq = b.req[req];
  const DevXillTable xt = b.xt[rq.prim];
  const double dd = rq.xdens_div;
  // XILL_BINS bins per thread, XILL_THREADS apart: the corner weights and rows
  // are set up once for all of them, and their loads are in flight together
  const int e0 = (blockIdx.x % chunks) * (XILL_THREADS * XILL_BINS) + threadIdx.x;
  if (e0 >= xt.n_ener) return;
  const ZoneRec zr = b.zone[rq.zone_off + wk.zone];
  int i0[4];
  double f[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {  // the zone's own Ecut and logXi brackets
    const bool ect = (k == xt.ect_dim), lxi = (k == xt.lxi_dim);
    i0[k] = ect ? zr.ect_ind : (lxi ? zr.lxi_ind : rq.xind[k]);
    f[k] = ect ? zr.ect_fac : (lxi ? zr.lxi_fac : rq.xfac[k]);
  }
  // corner c has the offsets of bits 3, 2, 1, 0 of kCorner[c] on axes 0..3
  constexpr unsigned char kCorner[16] = {0x0, 0x8, 0x4, 0x2, 0xC, 0xA, 0x6, 0xE,
                                         0x1, 0x9, 0x5, 0x3, 0xD, 0xB, 0x7, 0xF};
  double w[16];
  unsigned row[16];  // first row (inclination 0) of each corner
#pragma unroll
  for (int c = 0; c < 16; c++) {
    int o[4];
    double fk[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      o[k] = (kCorner[c] >> (3 - k)) & 1;
      fk[k] = o[k] ? f[k] : (1.0 - f[k]);
    }
    w[c] = fk[0] * fk[1] * fk[2] * fk[3];
    const unsigned r = (((unsigned)(i0[0] + o[0]) * xt.n_val[1] + (i0[1] + o[1])) * xt.n_val[2] + (i0[2] + o[2])) *
                           xt.n_val[3] + (i0[3] + o[3]);
    row[c] = r * xt.n_incl;
  }
  const double *dist = b.dist + rq.dist_off + (size_t)wk.zone * xt.n_incl;
  const size_t ne = (size_t)xt.n_ener;
  int eb[XILL_BINS];
  bool in[XILL_BINS];
#pragma unroll
  for (int u = 0; u < XILL_BINS; u++) {
    eb[u] = e0 + u * XILL_THREADS;
    in[u] = eb[u] < xt.n_ener;
    if (!in[u]) eb[u] = e0;  // a valid address; the result is not stored
  }
  double acc[XILL_BINS];
#pragma unroll
  for (int u = 0; u < XILL_BINS; u++) acc[u] = 0.0;
  const int n_o = xt.n_out > 1 ? 2 : 1;
  for (int a = 0; a < xt.n_incl; a++) {
    const double d = dist[a];
    if (d == 0.0) continue;
    double s[XILL_BINS];
#pragma unroll
    for (int u = 0; u < XILL_BINS; u++) s[u] = 0.0;
    for (int o = 0; o < n_o; o++) {
      // a six-parameter table: the five-parameter interpolation at both
      // brackets of its first axis, lerped (interp_6d_tab_incl,
      // src/xilltable.c:997-1020); five parameters: one pass with weight 1
      const double wo = n_o == 1 ? 1.0 : (o ? rq.xo_fac : 1.0 - rq.xo_fac);
      const size_t ob = (size_t)(rq.xo_ind + o) * xt.outer_rows;
      double so[XILL_BINS];
#pragma unroll
      for (int u = 0; u < XILL_BINS; u++) so[u] = 0.0;
#pragma unroll
      for (int c = 0; c < 16; c++) {
        const float *src = xt.spec + (ob + row[c] + a) * ne;
#pragma unroll
        for (int u = 0; u < XILL_BINS; u++) so[u] += w[c] * xval(src[eb[u]], dd);
      }
#pragma unroll
      for (int u = 0; u < XILL_BINS; u++) s[u] += wo * so[u];
    }
#pragma unroll
    for (int u = 0; u < XILL_BINS; u++) acc[u] += d * s[u];
  }
  double *dst = b.xspec + rq.xspec_off + (size_t)wk.zone * xt.n_ener;
#pragma unroll
  for (int u = 0; u < XILL_BINS; u++)
    if (in[u]) dst[eb[u]] = acc[u] / zr.ncf;
}

// relxilllp: the same for the zones of a lamp-post request (src/Relxill.cpp:
// 357-388), which differ only in the Ecut they see and in their angle
// distributions.  One unit per (request, slice of XILL_THREADS * XILL_BINS_Z
// table bins) for all zones, so that a table row is read once however many
// zones use it:
//   zone z = sum_a d_z[a] sum_c w_c row_c(a)
//          = sum_{(e, a)} k_z(e, a) sum_{c8} w_c8 row_{c8, e}(a),
//   k_z(e, a) = d_z[a] (1 - f_z) for e = i_z, d_z[a] f_z for e = i_z + 1, else 0,
// with (i_z, f_z) the zone's bracket on the Ecut axis and c8 the eight
// corners of the other three axes (their weights as the products of the
// reference's factors).  The block first lists the pairs (e, a) that some
// zone weights, with every zone's k_z, in LDS (in windows of XILL_THREADS
// candidates, in a fixed order), then streams the 8 rows of each listed pair.
// Tables without an Ecut axis use the request's bracket on the fourth axis
// for every zone.
template <int NZ>
__global__ __launch_bounds__(XILL_THREADS, 3) void k_xillz(BatchDev b, const XillWork *__restrict__ work, int chunks) {
  constexpr int NW = XILL_THREADS / 64;
  __shared__ int s_ind[NZ];
  __shared__ double s_fac[NZ];
  __shared__ unsigned s_off[XILL_THREADS];
  __shared__ double s_k[XILL_THREADS][NZ];
  __shared__ int s_wn[NW];
  const XillWork wk = work[blockIdx.x / chunks];
  const int req = wk.req;
  if (b.status[req] != 0) return;  // uniform over the block
  const ReqParams &rq = b.req[req];
  const DevXillTable xt = b.xt[rq.prim];
  const double dd = rq.xdens_div;
  const int nz = min(rq.n_zones, NZ);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int zax = xt.ect_dim >= 0 ? xt.ect_dim : 3;
  if (tid < nz) {
    const ZoneRec &zr = b.zone[rq.zone_off + tid];
    s_ind[tid] = xt.ect_dim >= 0 ? zr.ect_ind : rq.xind[3];
    s_fac[tid] = xt.ect_dim >= 0 ? zr.ect_fac : rq.xfac[3];
  }
  __syncthreads();
  int elo = s_ind[0], ehi = s_ind[0];
  for (int z = 1; z < nz; z++) {
    elo = min(elo, s_ind[z]);
    ehi = max(ehi, s_ind[z]);
  }
  const int ni = xt.n_incl;
  const int n_o = xt.n_out > 1 ? 2 : 1;
  const int n_cand = (ehi - elo + 2) * ni;  // candidate pairs (e, a), e in [elo, ehi + 1]
  // corner rows of the axes other than zax (zax index 0), in rows of ni
  // inclinations; one step on zax moves a row by `stride` rows
  double w8[8];
  unsigned cb[8];
  unsigned stride = 1;
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (k > zax) stride *= (unsigned)xt.n_val[k];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    double w = 1.0;
    unsigned r = 0;
    int bit = 2;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      int idx = 0;
      if (k != zax) {
        const int o = (c >> bit) & 1;
        bit--;
        w *= o ? rq.xfac[k] : (1.0 - rq.xfac[k]);
        idx = rq.xind[k] + o;
      }
      r = r * (unsigned)xt.n_val[k] + (unsigned)idx;
    }
    w8[c] = w;
    cb[c] = r * (unsigned)ni;
  }
  const double *dist = b.dist + rq.dist_off;
  const size_t ne = (size_t)xt.n_ener;
  const int e0 = (blockIdx.x % chunks) * (XILL_THREADS * XILL_BINS_Z) + tid;
  int eb[XILL_BINS_Z];
  bool in[XILL_BINS_Z];
#pragma unroll
  for (int u = 0; u < XILL_BINS_Z; u++) {
    eb[u] = e0 + u * XILL_THREADS;
    in[u] = eb[u] < xt.n_ener;
    if (!in[u]) eb[u] = 0;  // a valid address; the result is not stored
  }
  double acc[NZ][XILL_BINS_Z];
#pragma unroll
  for (int z = 0; z < NZ; z++)
#pragma unroll
    for (int u = 0; u < XILL_BINS_Z; u++) acc[z][u] = 0.0;
  for (int q0 = 0; q0 < n_cand; q0 += XILL_THREADS) {
    // list this window's used pairs: thread tid tests candidate q0 + tid
    const int q = q0 + tid;
    const int e = elo + q / ni, a = q % ni;
    auto kz = [&](int z) {  // k_z(e, a); zones beyond nz weigh nothing
      if (z >= nz) return 0.0;
      const double d = dist[z * ni + a];
      const int iz = s_ind[z];
      return e == iz ? d * (1.0 - s_fac[z]) : (e == iz + 1 ? d * s_fac[z] : 0.0);
    };
    bool used = false;
    if (q < n_cand)
      for (int z = 0; z < nz; z++) used |= kz(z) != 0.0;
    const unsigned long long m = __ballot(used);
    if (lane == 0) s_wn[wv] = __popcll(m);
    __syncthreads();
    int slot = __popcll(m & ((1ull << lane) - 1ull)), n_used = 0;
    for (int w = 0; w < NW; w++) {
      if (w < wv) slot += s_wn[w];
      n_used += s_wn[w];
    }
    if (used) {
      s_off[slot] = (unsigned)e * stride * (unsigned)ni + (unsigned)a;
      for (int z = 0; z < NZ; z++) s_k[slot][z] = kz(z);
    }
    __syncthreads();
    if (e0 < xt.n_ener) {
      for (int p = 0; p < n_used; p++) {
        const unsigned po = s_off[p];
        double s8[XILL_BINS_Z];
#pragma unroll
        for (int u = 0; u < XILL_BINS_Z; u++) s8[u] = 0.0;
        for (int o = 0; o < n_o; o++) {  // (a six-parameter table: both brackets of its first axis)
          const double wo = n_o == 1 ? 1.0 : (o ? rq.xo_fac : 1.0 - rq.xo_fac);
          const size_t ob = (size_t)(rq.xo_ind + o) * xt.outer_rows;
          double so[XILL_BINS_Z];
#pragma unroll
          for (int u = 0; u < XILL_BINS_Z; u++) so[u] = 0.0;
#pragma unroll
          for (int c = 0; c < 8; c++) {
            const float *src = xt.spec + (ob + cb[c] + po) * ne;
#pragma unroll
            for (int u = 0; u < XILL_BINS_Z; u++) so[u] += w8[c] * xval(src[eb[u]], dd);
          }
#pragma unroll
          for (int u = 0; u < XILL_BINS_Z; u++) s8[u] += wo * so[u];
        }
#pragma unroll
        for (int z = 0; z < NZ; z++) {
          const double k = s_k[p][z];
#pragma unroll
          for (int u = 0; u < XILL_BINS_Z; u++) acc[z][u] += k * s8[u];
        }
      }
    }
    __syncthreads();  // the list is rewritten by the next window
  }
  if (e0 >= xt.n_ener) return;
#pragma unroll
  for (int z = 0; z < NZ; z++) {
    if (z < nz) {
      const double ncf = b.zone[rq.zone_off + z].ncf;
      double *dst = b.xspec + rq.xspec_off + (size_t)z * ne;
#pragma unroll
      for (int u = 0; u < XILL_BINS_Z; u++)
        if (in[u]) dst[eb[u]] = acc[z][u] / ncf;
    }
  }
}

// Resident reflection set: per angle, rebin onto the convolution grid, band
// sum, and the real FFT of (spectrum * E/dE), stored for k = 0..n/2.
__global__ __launch_bounds__(CONV_THREADS_MAX) void k_refl_prepare(DevConvGrid cg, const double *refl_ener,
                                                                  const double *refl, int n_refl,
                                                                  double *spec, double *bandsum, double *fft) {
  extern __shared__ double2 s_buf[];
  __shared__ double s_w[CONV_THREADS_MAX / 64];
  const int a = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, nwave = blockDim.x >> 6;
  const int n = cg.n, m = n >> 1;
  const double *f0 = refl + (size_t)a * n_refl;
  const GridGuess gi = grid_guess(refl_ener, n_refl);
  auto ge = [&](int k) { return refl_ener[k]; };
  const int nb = n / nwave;
  double sx = 0.0;
  for (int q = 0; q < nb; q += 64) {
    const int i = wave * nb + q + lane;
    const double elo = cg.ener[i], ehi = cg.ener[i + 1];
    const int clo = count_le(ge, n_refl, elo, gi);
    const int chi = count_le(ge, n_refl, ehi, gi);
    const double xr = rebin_bin(elo, ehi, clo, chi, ge, n_refl, [&](int k) { return f0[k]; });
    spec[(size_t)a * n + i] = xr;
    if (elo >= 0.01 && ehi < 1000.0) sx += xr;
    s_buf[P(i)] = make_double2(xr * cg.cfac[i], 0.0);
  }
  const double s = block_sum4(sx, s_w);
  if (tid == 0) bandsum[a] = s;
  lds_fft<-1>(s_buf, n, reinterpret_cast<const double2 *>(cg.twp_fwd));
  for (int k = tid; k <= m; k += blockDim.x) {
    fft[((size_t)a * (m + 1) + k) * 2 + 0] = s_buf[P(k)].x;
    fft[((size_t)a * (m + 1) + k) * 2 + 1] = s_buf[P(k)].y;
  }
}

// plain streaming copy, used to calibrate the HBM byte counters
__global__ __launch_bounds__(256) void k_copy(const double4 *__restrict__ src, double4 *__restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    dst[i] = src[i];
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int launch_setup(const BatchDev &b, void *stream) {
  if (b.n_req <= 0) return 0;
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, (hipStream_t)stream, b.ctr);
  if (b.max_zones > LP_ZONES)
    hipLaunchKernelGGL(k_setup<MAX_ZONES>, dim3(b.n_req), dim3(SETUP_THREADS), 0, (hipStream_t)stream, b);
  else
    hipLaunchKernelGGL(k_setup<LP_ZONES>, dim3(b.n_req), dim3(SETUP_THREADS), 0, (hipStream_t)stream, b);
  return (int)hipGetLastError();
}

static size_t row_lds(const BatchDev &b, bool cos_rows, size_t elem) {
  return (size_t)(cos_rows ? 2 : 1) * 64 * (2 * b.rt.n_g + 1) * elem;
}

// LDS of one k_integrate wavefront: lane rows, then (FP64 path) 64 weight
// slots, then the SUM_BINS x SUM_STRIDE parked values, which double as the
// node-row staging area.  *stage_rows: node rows the staging area takes, as
// many as fit without lowering the wavefronts per CU that the LDS allows (at
// most 12, the VGPR limit of the FP32 path), up to 8 and at least 2.
static size_t integrate_lds(const BatchDev &b, bool cos_rows, size_t elem, int *stage_rows) {
  const size_t rows = integrate_rows_bytes(b.rt.n_g, cos_rows, elem);
  const size_t w = (elem == sizeof(double)) ? 64 * sizeof(double) : 0;
  const size_t park = (size_t)SUM_BINS * SUM_STRIDE * elem;
  const size_t nq = 2 * (size_t)b.rt.n_g;
  // wavefronts per CU that the LDS holds: gfx950 allocates a workgroup's LDS
  // in 1280-byte granules, 128 per CU (160 KB); at most 16 (the VGPRs of the
  // FP32 forms allow four per SIMD)
  auto per_cu = [](size_t lds) { return std::min<size_t>(16, 128 / ((lds + 1279) / 1280)); };
  const size_t base = rows + w + park;
  for (int c = 8; c > 2; c--) {
    if ((size_t)c * nq > (size_t)STAGE_MAX) continue;
    const size_t need = rows + w + std::max(park, (size_t)c * nq * sizeof(double));
    if (per_cu(need) >= per_cu(base)) {
      *stage_rows = c;
      return need;
    }
  }
  *stage_rows = 2;  // the two rows a radius lerps between (2 nq <= STAGE_MAX: n_g <= MAX_NG)
  return rows + w + std::max(park, (size_t)2 * nq * sizeof(double));
}

int launch_integrate(const BatchDev &b, int n_blocks, bool cos_rows, void *stream) {
  if (b.n_req <= 0 || n_blocks <= 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  int stage = 0;
  const size_t lds = integrate_lds(b, cos_rows, b.mode == 0 ? sizeof(double) : sizeof(float), &stage);
  auto go = [&](auto kern) {
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(kern, dim3(n_blocks), dim3(64), lds, st, b, stage);
  };
  if (b.mode == 0) {
    if (cos_rows) go(k_integrate<0, 1>);
    else go(k_integrate<0, 0>);
  } else {
    if (cos_rows) go(k_integrate<1, 1>);
    else go(k_integrate<1, 0>);
  }
  return (int)hipGetLastError();
}

int launch_fixup(const BatchDev &b, bool cos_rows, void *stream) {
  if (b.mode != 1 || b.n_req <= 0) return 0;
  const size_t lds = row_lds(b, cos_rows, sizeof(double));
  (void)hipFuncSetAttribute((const void *)k_fixup, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k_fixup, dim3(1024), dim3(64), lds, (hipStream_t)stream, b, (int)cos_rows);
  return (int)hipGetLastError();
}

int launch_finalize(const BatchDev &b, void *stream) {
  if (b.n_req <= 0) return 0;
  hipLaunchKernelGGL(k_finalize, dim3(b.n_req), dim3(SETUP_THREADS), 0, (hipStream_t)stream, b);
  return (int)hipGetLastError();
}

static size_t conv_lds(int n) { return (size_t)(n + n / 16) * sizeof(double2); }
static int conv_threads(int n) { return std::max(CONV_THREADS, n / 8); }

int launch_conv(const BatchDev &b, const XillWork *conv_work, int n_conv_work, void *stream) {
  if (n_conv_work <= 0) return 0;
  const size_t lds = conv_lds(b.cg.n);
  (void)hipFuncSetAttribute((const void *)k_conv, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k_conv, dim3(n_conv_work), dim3(conv_threads(b.cg.n)), lds, (hipStream_t)stream, b, conv_work);
  return (int)hipGetLastError();
}

// the larger energy grid of the two xillver tables (units beyond a request's
// own table return at once)
static int xill_n_ener(const BatchDev &b) {
  int n = 0;
  for (int pr = 0; pr < N_PRIM; pr++) n = std::max(n, b.xt[pr].n_ener);
  return n;
}

int launch_xill(const BatchDev &b, const XillWork *work, int n_single, int n_multi, void *stream) {
  if (n_single > 0) {
    const int chunks = (xill_n_ener(b) + XILL_THREADS * XILL_BINS - 1) / (XILL_THREADS * XILL_BINS);
    hipLaunchKernelGGL(k_xill1, dim3(n_single * chunks), dim3(XILL_THREADS), 0, (hipStream_t)stream, b,
                       work, chunks);
  }
  if (n_multi > 0) {
    const int chunks = (xill_n_ener(b) + XILL_THREADS * XILL_BINS_Z - 1) / (XILL_THREADS * XILL_BINS_Z);
    hipLaunchKernelGGL(k_xillz<LP_ZONES>, dim3(n_multi * chunks), dim3(XILL_THREADS), 0,
                       (hipStream_t)stream, b, work + n_single, chunks);
  }
  return (int)hipGetLastError();
}

int launch_rrad_corr(const BatchDev &b, const XillWork *work, int n_work, void *stream) {
  if (n_work <= 0) return 0;
  const size_t lds = sizeof(double) * (size_t)xill_n_ener(b);
  (void)hipFuncSetAttribute((const void *)k_rrad_corr, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k_rrad_corr, dim3(n_work), dim3(RRC_THREADS), lds, (hipStream_t)stream, b, work);
  return (int)hipGetLastError();
}

int launch_nth(const BatchDev &b, const XillWork *work, int n_work, void *stream) {
  if (n_work <= 0) return 0;
  hipLaunchKernelGGL(k_nth, dim3(n_work), dim3(NTH_THREADS_R), 0, (hipStream_t)stream, b, work);
  return (int)hipGetLastError();
}

int launch_primary(const BatchDev &b, const int *reqs, int n_reqs, void *stream) {
  if (n_reqs <= 0) return 0;
  hipLaunchKernelGGL(k_primary, dim3(n_reqs), dim3(PRIM_THREADS), 0, (hipStream_t)stream, b, reqs);
  return (int)hipGetLastError();
}

int launch_copy(const void *src, void *dst, size_t bytes, void *stream) {
  const size_t n = bytes / sizeof(double4);
  hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, (hipStream_t)stream, (const double4 *)src, (double4 *)dst, n);
  return (int)hipGetLastError();
}

int launch_refl_prepare(const DevConvGrid &cg, const double *refl_ener, const double *refl, int n_refl,
                        int n_incl, double *spec, double *bandsum, double *fft, void *stream) {
  const size_t lds = conv_lds(cg.n);
  (void)hipFuncSetAttribute((const void *)k_refl_prepare, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k_refl_prepare, dim3(n_incl), dim3(conv_threads(cg.n)), lds, (hipStream_t)stream, cg, refl_ener,
                     refl, n_refl, spec, bandsum, fft);
  return (int)hipGetLastError();
}

}  // namespace rx
