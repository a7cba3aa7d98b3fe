// rx_math.h -- per-element math of the relline path, shared by the HIP kernels
// (compiled for gfx950) and by the host-compiled kernel-emulation used in CPU
// unit tests of these building blocks.
//
// Reference correspondence (relxill 2.4):
//   g* grid ................. src/Relprofile.cpp:104-116
//   relb_func ............... src/Relprofile.cpp:489-521
//   romberg_integration ..... src/Relprofile.cpp:524-579
//   int_edge ................ src/Relprofile.cpp:585-621
//   int_romb ................ src/Relprofile.cpp:628-647
//   integ_relline_bin ....... src/Relprofile.cpp:650-726
// Restructuring (same results up to rounding):
//   * the integrand is evaluated for both transfer-function branches k=0,1 at
//     once (the reference calls relb_func per k at identical points);
//   * Romberg re-uses the previous level's points (the reference re-evaluates
//     every interior point at every level) -- the points are bit-identical
//     (a + pas*i with pas halved exactly), only the summation order differs;
//   * the per-radius factor emis/(gmax-gmin) of relb_func is applied once per
//     bin integral instead of once per evaluation;
//   * the g* bracket is found by a direct index computation corrected to the
//     exact binary_search result (src/relutility.c:155-171).
// FP32 mode keeps every decision in FP64 (bin -> g* mapping, edge tests, g*
// bracket, red-wing switch) and evaluates the integrand and the Romberg
// tables in FP32; a Romberg stop test whose FP32 precision estimate falls
// within a guard band of the 2% threshold is recomputed in FP64, so both modes
// stop at the level the FP64 algorithm stops.  The 1/sqrt endpoint integrals
// (int_edge) are FP64 in both modes.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RX_HD __host__ __device__
#else
#include <cmath>
#define RX_HD
#endif

#include <math.h>
#include <stdint.h>

namespace rx {

constexpr double GFAC_H = 5e-3;          // src/Relprofile.cpp:36
constexpr double ROMB_PREC = 0.02;       // src/Relprofile.cpp:525
constexpr int ROMB_ITERMAX = 5;          // src/Relprofile.cpp:528 (niter <= 6)
constexpr double RED_WING = 0.95;        // src/Relprofile.cpp:636
constexpr float ROMB_GUARD = 2e-3f;      // relative guard band around ROMB_PREC (FP32 mode)
constexpr double PI = 3.14159265358979323846;

// Per-radius view used by the bin integral.  T64/C64 hold the fine-radius
// transfer function / emission-angle cosines [ng][2] in FP64, T32/C32 the same
// rounded to FP32 (FP32 mode only).
struct RadiusView {
  double gmin, gmax, del_g;
  const double *T64;
  const double *C64;
  const float *T32;
  const float *C32;
  int ng;
  int limb;  // 0 none, 1 Laor, 2 Haardt
  double gs_step;   // (1-2H)/(ng-1)
  double inv_step;  // 1/gs_step
};

RX_HD inline double gstar_node(double step, int i) {
#pragma clang fp contract(off)
  return GFAC_H + step * (double)((float)i);
}

// Index k in [0, ng-2] with gstar[k] <= x < gstar[k+1], clamped as
// binary_search does (0 below the grid, ng-2 above it).
RX_HD inline int g_bracket(double x, int ng, double step, double inv_step) {
#pragma clang fp contract(off)
  double q = (x - GFAC_H) * inv_step;
  int i;
  if (!(q > 0.0)) {
    i = 0;
  } else if (q >= (double)(ng - 2)) {
    i = ng - 2;
  } else {
    i = (int)q;
  }
  if (i < ng - 2 && gstar_node(step, i + 1) <= x) i++;
  if (i > 0 && gstar_node(step, i) > x) i--;
  return i;
}

RX_HD inline double limb_factor64(int limb, double fmu0) {
  if (limb == 1) return 1.0 + 2.06 * fmu0;       // Laor (1991)
  if (limb == 2) return log(1.0 + 1.0 / fmu0);   // Haardt (1993)
  return 1.0;
}

RX_HD inline float limb_factor32(int limb, float fmu0) {
  if (limb == 1) return 1.0f + 2.06f * fmu0;
  if (limb == 2) return logf(1.0f + 1.0f / fmu0);
  return 1.0f;
}

// Reduced integrand eg^3 ftrf_k / sqrt(g*(1-g*)) [* limb] for k = 0 and 1;
// relb_func = reduced * emis / (gmax - gmin).
RX_HD inline void eval64_pair(double eg, const RadiusView &v, double &f0, double &f1) {
#pragma clang fp contract(off)
  const double egs = (eg - v.gmin) * v.del_g;
  const int ind = g_bracket(egs, v.ng, v.gs_step, v.inv_step);
  const double g0 = gstar_node(v.gs_step, ind);
  const double g1 = gstar_node(v.gs_step, ind + 1);
  const double inte = (egs - g0) / (g1 - g0);
  const double inte1 = 1.0 - inte;
  const double base = eg * eg * eg / sqrt(egs - egs * egs);
  const double t0 = inte * v.T64[ind * 2 + 0] + inte1 * v.T64[(ind + 1) * 2 + 0];
  const double t1 = inte * v.T64[ind * 2 + 1] + inte1 * v.T64[(ind + 1) * 2 + 1];
  f0 = base * t0;
  f1 = base * t1;
  if (v.limb) {
    const double m0 = inte * v.C64[ind * 2 + 0] + inte1 * v.C64[(ind + 1) * 2 + 0];
    const double m1 = inte * v.C64[ind * 2 + 1] + inte1 * v.C64[(ind + 1) * 2 + 1];
    f0 *= limb_factor64(v.limb, m0);
    f1 *= limb_factor64(v.limb, m1);
  }
}

RX_HD inline float rx_rsqrtf(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / sqrtf(x);
#endif
}

// FP32 integrand: coordinates and bracket in FP64, values in FP32.
RX_HD inline void eval32_pair(double eg, const RadiusView &v, float &f0, float &f1) {
  const double egs = (eg - v.gmin) * v.del_g;
  const int ind = g_bracket(egs, v.ng, v.gs_step, v.inv_step);
  const double g0 = gstar_node(v.gs_step, ind);
  const float inte = (float)((egs - g0) * v.inv_step);
  const float t = (float)egs;
  const float u = (float)(1.0 - egs);
  const float e = (float)eg;
  const float base = e * e * e * rx_rsqrtf(t * u);
  const float a0 = v.T32[ind * 2 + 0], b0 = v.T32[(ind + 1) * 2 + 0];
  const float a1 = v.T32[ind * 2 + 1], b1 = v.T32[(ind + 1) * 2 + 1];
  f0 = base * fmaf(inte, a0 - b0, b0);
  f1 = base * fmaf(inte, a1 - b1, b1);
  if (v.limb) {
    const float c0 = v.C32[ind * 2 + 0], d0 = v.C32[(ind + 1) * 2 + 0];
    const float c1 = v.C32[ind * 2 + 1], d1 = v.C32[(ind + 1) * 2 + 1];
    f0 *= limb_factor32(v.limb, fmaf(inte, c0 - d0, d0));
    f1 *= limb_factor32(v.limb, fmaf(inte, c1 - d1, d1));
  }
}

// Romberg for both branches on [a,b], FP64.  Returns t_0[niter0][0] + t_1[niter1][0].
RX_HD inline double romberg64_pair(double a, double b, const RadiusView &v, int *levels = nullptr) {
#pragma clang fp contract(off)
  double fa0, fa1, fb0, fb1;
  eval64_pair(a, v, fa0, fa1);
  eval64_pair(b, v, fb0, fb1);
  const double ta0 = (fa0 + fb0) / 2.0, ta1 = (fa1 + fb1) / 2.0;
  double pas = b - a;
  double s0 = ta0, s1 = ta1;  // ta + sum of interior points so far
  double p0[ROMB_ITERMAX + 2], p1[ROMB_ITERMAX + 2];  // previous anti-diagonal
  double c0[ROMB_ITERMAX + 2], c1[ROMB_ITERMAX + 2];
  p0[0] = ta0 * pas;
  p1[0] = ta1 * pas;
  double res0 = p0[0], res1 = p1[0];
  bool run0 = true, run1 = true;
  int lev = 0;
#pragma unroll
  for (int n = 1; n <= ROMB_ITERMAX + 1; n++) {
    if (!(run0 || run1)) break;
    lev = n;
    pas = pas / 2.0;
    const int nnew = 1 << (n - 1);
    double add0 = 0.0, add1 = 0.0;
    for (int m = 0; m < nnew; m++) {
      double g0v, g1v;
      eval64_pair(a + pas * (double)(2 * m + 1), v, g0v, g1v);
      add0 += g0v;
      add1 += g1v;
    }
    s0 += add0;
    s1 += add1;
    c0[0] = s0 * pas;
    c1[0] = s1 * pas;
    double r = 1.0;
#pragma unroll
    for (int ii = 1; ii <= ROMB_ITERMAX + 1; ii++) {
      if (ii > n) break;
      r *= 4.0;
      c0[ii] = (r * c0[ii - 1] - p0[ii - 1]) / (r - 1.0);
      c1[ii] = (r * c1[ii - 1] - p1[ii - 1]) / (r - 1.0);
    }
    if (run0) {
      res0 = c0[n];
      const double ob = fabs(c0[n] - p0[n - 1]) / c0[n];
      run0 = (ob > ROMB_PREC) && (n <= ROMB_ITERMAX);
    }
    if (run1) {
      res1 = c1[n];
      const double ob = fabs(c1[n] - p1[n - 1]) / c1[n];
      run1 = (ob > ROMB_PREC) && (n <= ROMB_ITERMAX);
    }
#pragma unroll
    for (int ii = 0; ii <= ROMB_ITERMAX + 1; ii++) {
      if (ii > n) break;
      p0[ii] = c0[ii];
      p1[ii] = c1[ii];
    }
  }
  if (levels) *levels = lev;
  return res0 + res1;
}

// Romberg for both branches, FP32 values.  Sets *need64 when a stop decision
// fell inside the guard band; the caller then recomputes in FP64.
RX_HD inline float romberg32_pair(double a, double b, const RadiusView &v, bool *need64) {
  float fa0, fa1, fb0, fb1;
  eval32_pair(a, v, fa0, fa1);
  eval32_pair(b, v, fb0, fb1);
  const float ta0 = (fa0 + fb0) * 0.5f, ta1 = (fa1 + fb1) * 0.5f;
  double pas = b - a;
  float s0 = ta0, s1 = ta1;
  float p0[ROMB_ITERMAX + 2], p1[ROMB_ITERMAX + 2];
  float c0[ROMB_ITERMAX + 2], c1[ROMB_ITERMAX + 2];
  const float pf = (float)pas;
  p0[0] = ta0 * pf;
  p1[0] = ta1 * pf;
  float res0 = p0[0], res1 = p1[0];
  bool run0 = true, run1 = true;
  const float lo = (float)ROMB_PREC * (1.0f - ROMB_GUARD);
  const float hi = (float)ROMB_PREC * (1.0f + ROMB_GUARD);
#pragma unroll
  for (int n = 1; n <= ROMB_ITERMAX + 1; n++) {
    if (!(run0 || run1)) break;
    pas = pas * 0.5;
    const int nnew = 1 << (n - 1);
    float add0 = 0.0f, add1 = 0.0f;
    for (int m = 0; m < nnew; m++) {
      float g0v, g1v;
      eval32_pair(a + pas * (double)(2 * m + 1), v, g0v, g1v);
      add0 += g0v;
      add1 += g1v;
    }
    s0 += add0;
    s1 += add1;
    const float pfn = (float)pas;
    c0[0] = s0 * pfn;
    c1[0] = s1 * pfn;
    float r = 1.0f;
#pragma unroll
    for (int ii = 1; ii <= ROMB_ITERMAX + 1; ii++) {
      if (ii > n) break;
      r *= 4.0f;
      const float inv = 1.0f / (r - 1.0f);
      c0[ii] = (r * c0[ii - 1] - p0[ii - 1]) * inv;
      c1[ii] = (r * c1[ii - 1] - p1[ii - 1]) * inv;
    }
    if (run0) {
      res0 = c0[n];
      const float ob = fabsf(c0[n] - p0[n - 1]) / c0[n];
      if (ob >= lo && ob <= hi && n <= ROMB_ITERMAX) *need64 = true;
      run0 = (ob > (float)ROMB_PREC) && (n <= ROMB_ITERMAX);
    }
    if (run1) {
      res1 = c1[n];
      const float ob = fabsf(c1[n] - p1[n - 1]) / c1[n];
      if (ob >= lo && ob <= hi && n <= ROMB_ITERMAX) *need64 = true;
      run1 = (ob > (float)ROMB_PREC) && (n <= ROMB_ITERMAX);
    }
#pragma unroll
    for (int ii = 0; ii <= ROMB_ITERMAX + 1; ii++) {
      if (ii > n) break;
      p0[ii] = c0[ii];
      p1[ii] = c1[ii];
    }
  }
  return res0 + res1;
}

// int_edge in reduced units (the trailing (gmax-gmin) cancels the Jacobian
// inside relb_func), FP64 in both modes.
RX_HD inline double int_edge64(double blo, double bhi, const RadiusView &v) {
#pragma clang fp contract(off)
  double hex, lo, hi;
  if (blo <= 0.5) {
    hex = GFAC_H;
    lo = blo;
    hi = bhi;
  } else {
    hex = 1.0 - GFAC_H;
    lo = 1.0 - bhi;
    hi = 1.0 - blo;
  }
  const double e = (hex * (v.gmax - v.gmin) + v.gmin) * 1.0;
  double f0, f1;
  eval64_pair(e, v, f0, f1);
  double norm = 0.0;
  norm = norm + f0;
  norm = norm + f1;
  norm = norm * sqrt(GFAC_H);
  return 2 * norm * (sqrt(hi) - sqrt(lo)) * 1.0 * (v.gmax - v.gmin);
}

// Integral of one energy bin [rlo0, rhi0] (line-energy units) at one radius,
// in reduced units (multiply by emis/(gmax-gmin) for relb_func units).
// MODE 0: FP64; MODE 1: FP32 interior with FP64 decisions and edges.
template <int MODE>
RX_HD inline double integ_bin(double rlo0, double rhi0, const RadiusView &v) {
#pragma clang fp contract(off)
  double gblo = (rlo0 / 1.0 - v.gmin) * v.del_g;
  if (gblo < 0.0) gblo = 0.0; else if (gblo > 1.0) gblo = 1.0;
  double gbhi = (rhi0 / 1.0 - v.gmin) * v.del_g;
  if (gbhi < 0.0) gbhi = 0.0; else if (gbhi > 1.0) gbhi = 1.0;
  if (gbhi == 0) return 0.0;
  double rlo = rlo0, rhi = rhi0;
  double flu = 0.0;
  if (gblo <= GFAC_H) {
    double hlo = gblo, hhi = GFAC_H;
    rlo = (GFAC_H * (v.gmax - v.gmin) + v.gmin) * 1.0;
    if (gbhi <= GFAC_H) {
      hhi = gbhi;
      rlo = -1.0;
    }
    flu = flu + int_edge64(hlo, hhi, v);
  }
  if (gbhi >= (1.0 - GFAC_H)) {
    double hhi = gbhi, hlo = 1.0 - GFAC_H;
    rhi = ((1 - GFAC_H) * (v.gmax - v.gmin) + v.gmin) * 1.0;
    if (gblo >= (1.0 - GFAC_H)) {
      hlo = gblo;
      rhi = -1.0;
    }
    flu = flu + int_edge64(hlo, hhi, v);
  }
  if ((rhi >= 0) && (rlo >= 0)) {
    if (rlo >= 1.0 * RED_WING) {
      if (MODE == 0) {
        flu = flu + romberg64_pair(rlo, rhi, v);
      } else {
        bool need64 = false;
        const float r32 = romberg32_pair(rlo, rhi, v, &need64);
        flu = flu + (need64 ? romberg64_pair(rlo, rhi, v) : (double)r32);
      }
    } else {
      const double mid = (rhi + rlo) / 2.0;
      if (MODE == 0) {
        double f0, f1;
        eval64_pair(mid, v, f0, f1);
        flu = flu + (f0 * (rhi - rlo) + f1 * (rhi - rlo));
      } else {
        float f0, f1;
        eval32_pair(mid, v, f0, f1);
        flu = flu + (double)((f0 + f1) * (float)(rhi - rlo));
      }
    }
  }
  return flu;
}

// ---- small helpers shared by setup code ---------------------------------

RX_HD inline double kerr_rms_hd(double a) {
  double sgn = (a < 0) ? -1.0 : 1.0;
  double z1 = 1.0 + pow(1.0 - a * a, 1.0 / 3.0) * (pow(1.0 + a, 1.0 / 3.0) + pow(1.0 - a, 1.0 / 3.0));
  double z2 = sqrt((3.0 * a * a) + (z1 * z1));
  return 3.0 + z2 - sgn * sqrt((3.0 - z1) * (3.0 + z1 + (2 * z2)));
}

RX_HD inline double lerp1(double ifac, double lo, double hi) {
#pragma clang fp contract(off)
  return ifac * hi + (1.0 - ifac) * lo;  // interp_lin_1d, src/relutility.c:24-26
}

RX_HD inline double lerp2f(double i1, double i2, float r11, float r12, float r21, float r22) {
#pragma clang fp contract(off)
  // interp_lin_2d_float, src/relutility.c:53-58
  return (1.0 - i1) * (1.0 - i2) * r11 + (i1) * (1.0 - i2) * r12 + (1.0 - i1) * (i2) * r21 +
         (i1) * (i2) * r22;
}

// ascending search: k with arr[k] <= val < arr[k+1] (src/relutility.c:155-171)
template <class T, class V>
RX_HD inline int bsearch_asc(const T *arr, int n, V val) {
  if (n <= 1) return -1;
  int klo = 0, khi = n - 1;
  while ((khi - klo) > 1) {
    int k = (khi + klo) / 2;
    if (arr[k] > val) khi = k; else klo = k;
  }
  return klo;
}

// descending search (src/relutility.c:195-211)
template <class T, class V>
RX_HD inline int bsearch_desc(const T *arr, int n, V val) {
  if (n <= 1) return -1;
  int klo = 0, khi = n - 1;
  while ((khi - klo) > 1) {
    int k = (khi + klo) / 2;
    if (arr[k] < val) khi = k; else klo = k;
  }
  return klo;
}

// fine radial grid point ii (src/relutility.c:666-677)
RX_HD inline double fine_radius(double rin, double rout, int ii, int nr) {
#pragma clang fp contract(off)
  const double r1 = 1.0 / sqrt(rout);
  const double r2 = 1.0 / sqrt(rin);
  double x = ((double)(ii)) * (r2 - r1) / (nr - 1) + r1;
  const double y = 1.0 / x;
  return y * y;
}

// half of 2 pi r dr of a descending grid (src/relutility.c:233-244)
RX_HD inline double trapez_half(double rprev, double r, double rnext, int ii, int nr) {
#pragma clang fp contract(off)
  double dr;
  if (ii == 0) {
    dr = 0.5 * (r - rnext);
  } else if (ii == nr - 1) {
    dr = 0.5 * (rprev - r);
  } else {
    dr = 0.5 * (rprev - rnext);
  }
  return r * dr * PI;
}

}  // namespace rx
