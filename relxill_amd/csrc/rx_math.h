// rx_math.h -- per-element math of the relline path, shared by the HIP kernels
// (compiled for gfx950) and by host builds of the same functions.
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
//   * the value of the integrand at g* = H and 1-H that int_edge needs is a
//     per-radius constant, computed once per radius (edge norms);
//   * the g* bracket is found by a direct index computation corrected to the
//     exact binary_search result (src/relutility.c:155-171).
// FP32 mode keeps every decision in FP64 (bin -> g* mapping, edge tests,
// red-wing switch) and evaluates the integrand in bin-local FP32 coordinates:
// g*(s) = g*_lo + s dg for s in [0,1].  A point whose FP32 g* lies within a few
// rounding errors of a g*-grid node is re-done with FP64 coordinates, so the
// g* bracket always equals the FP64 one.  The Romberg tables run in FP32; a stop
// test whose FP32 estimate falls within a guard band of the 2% threshold is
// flagged and re-done in FP64 by the caller, so both modes stop at the level
// the FP64 algorithm stops.  The edge integrals' bounds stay FP64 in both
// modes (FP32 mode evaluates them in the cancellation-free quotient form).
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

#include <type_traits>

namespace rx {

constexpr double GFAC_H = 5e-3;          // src/Relprofile.cpp:36
constexpr double ROMB_PREC = 0.02;       // src/Relprofile.cpp:525
constexpr int ROMB_ITERMAX = 5;          // src/Relprofile.cpp:528 (niter <= 6)
constexpr double RED_WING = 0.95;        // src/Relprofile.cpp:636
// default relative guard band of the FP32 stop tests: the FP32 estimate of
// |R - top| / R carries a relative error of about 1e-6 (a few roundings of
// R's magnitude), 5e-5 of the 2% threshold; the band is ten times that
constexpr float ROMB_GUARD = 5e-4f;
constexpr double PI = 3.14159265358979323846;

// Per-radius view used by the bin integral.
struct RadiusView {
  double gmin, gmax, del_g;
  double normL, normU;  // sqrt(H) * sum_k f_k at g* = H and g* = 1-H (reduced units)
  const double *T64;    // FP64 mode: transfer function [ng][2]
  const double *C64;    // FP64 mode: emission-angle cosines [ng][2] (limb != 0)
  const float *T32;     // FP32 mode
  const float *C32;
  int ng;
  int limb;  // 0 none, 1 Laor, 2 Haardt
  double gs_step;   // (1-2H)/(ng-1)
  double inv_step;  // 1/gs_step
  float guard;      // FP32 mode: relative guard band around the Romberg threshold
  float eL, eU;     // FP32 mode: 2 norm (gmax - gmin) of the lower / upper edge integral
};

RX_HD inline double lerp1_(double ifac, double lo, double hi) {
#pragma clang fp contract(off)
  return ifac * hi + (1.0 - ifac) * lo;  // interp_lin_1d, src/relutility.c:24-26
}

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

// ---------------------------------------------------------------------------
// FP64 integrand
// ---------------------------------------------------------------------------

// Reduced integrand eg^3 ftrf_k / sqrt(g*(1-g*)) [* limb] for k = 0 and 1 from
// t4 = {T[ind][0], T[ind][1], T[ind+1][0], T[ind+1][1]} (relb_func's swapped
// lerp weights, src/Relprofile.cpp:502-504); relb_func = reduced * emis/(gmax-gmin).
RX_HD inline void eval64_at(double eg, double egs, int ind, const RadiusView &v, const double *t4,
                            const double *c4, double &f0, double &f1) {
#pragma clang fp contract(off)
  const double g0 = gstar_node(v.gs_step, ind);
  const double g1 = gstar_node(v.gs_step, ind + 1);
  const double inte = (egs - g0) / (g1 - g0);
  const double inte1 = 1.0 - inte;
  const double base = eg * eg * eg / sqrt(egs - egs * egs);
  const double t0 = inte * t4[0] + inte1 * t4[2];
  const double t1 = inte * t4[1] + inte1 * t4[3];
  f0 = base * t0;
  f1 = base * t1;
  if (v.limb) {
    const double m0 = inte * c4[0] + inte1 * c4[2];
    const double m1 = inte * c4[1] + inte1 * c4[3];
    f0 *= limb_factor64(v.limb, m0);
    f1 *= limb_factor64(v.limb, m1);
  }
}

RX_HD inline void eval64_pair(double eg, const RadiusView &v, double &f0, double &f1) {
#pragma clang fp contract(off)
  const double egs = (eg - v.gmin) * v.del_g;
  const int ind = g_bracket(egs, v.ng, v.gs_step, v.inv_step);
  eval64_at(eg, egs, ind, v, v.T64 + ind * 2, v.limb ? v.C64 + ind * 2 : nullptr, f0, f1);
}

// Edge norm of int_edge (src/Relprofile.cpp:609-618): sqrt(H) * sum_k relb_func
// at g* = hex, in reduced units.  T64/C64 must hold the radius' FP64 values.
RX_HD inline double edge_norm64(bool upper, const RadiusView &v) {
#pragma clang fp contract(off)
  const double hex = upper ? 1.0 - GFAC_H : GFAC_H;
  const double e = (hex * (v.gmax - v.gmin) + v.gmin) * 1.0;
  double f0, f1;
  eval64_pair(e, v, f0, f1);
  double norm = 0.0;
  norm = norm + f0;
  norm = norm + f1;
  return norm * sqrt(GFAC_H);
}

// The point of int_edge's norm (src/Relprofile.cpp:609-618): the energy at
// g* = H (lower) or 1 - H (upper), its g* and its g* bracket.
struct EdgePt {
  double e, egs;
  int ind;
};
RX_HD inline EdgePt edge_point(bool upper, const RadiusView &v) {
#pragma clang fp contract(off)
  const double hex = upper ? 1.0 - GFAC_H : GFAC_H;
  EdgePt p;
  p.e = (hex * (v.gmax - v.gmin) + v.gmin) * 1.0;
  p.egs = (p.e - v.gmin) * v.del_g;
  p.ind = g_bracket(p.egs, v.ng, v.gs_step, v.inv_step);
  return p;
}

// edge_norm64 from the radius' FP64 values at the point's bracket: t4 =
// {T[ind][0], T[ind][1], T[ind+1][0], T[ind+1][1]}, c4 the same for the
// cosines (read when limb != 0).  The values are those of a lerped FP64 row,
// so the norm has the same bits as edge_norm64's.
RX_HD inline double edge_norm_at(const EdgePt &p, const RadiusView &v, const double *t4, const double *c4) {
#pragma clang fp contract(off)
  double f0, f1;
  eval64_at(p.e, p.egs, p.ind, v, t4, c4, f0, f1);
  double norm = 0.0;
  norm = norm + f0;
  norm = norm + f1;
  return norm * sqrt(GFAC_H);
}

// int_edge (src/Relprofile.cpp:585-621) from the precomputed edge norms.
RX_HD inline double int_edge_pre(double blo, double bhi, const RadiusView &v) {
#pragma clang fp contract(off)
  double lo, hi, norm;
  if (blo <= 0.5) {
    lo = blo;
    hi = bhi;
    norm = v.normL;
  } else {
    lo = 1.0 - bhi;
    hi = 1.0 - blo;
    norm = v.normU;
  }
  return 2 * norm * (sqrt(hi) - sqrt(lo)) * 1.0 * (v.gmax - v.gmin);
}

// Integrand values at a bin end, handed from one bin to the next: a bin's
// upper end is the next bin's lower end, so a Romberg bin following a Romberg
// bin reuses the pair instead of evaluating it again (same arguments, same
// bits).  ok: f0/f1 hold the pair at the current bin's lower edge.
template <class T>
struct EdgePair {
  T f0, f1;
  bool ok;
};

RX_HD inline float rx_rcpf(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}


// Romberg for both branches on [a,b], FP64.  Returns t_0[niter0][0] + t_1[niter1][0].
// R[ii] holds the current anti-diagonal t[ii][n-ii] of the reference's table
// (src/Relprofile.cpp:559-576), updated in place level by level.  With
// have_a, (fa0, fa1) is the pair at a; the pair at b is returned in (fb0, fb1).
RX_HD inline double romberg64_pair(double a, double b, const RadiusView &v, bool have_a, double fa0,
                                   double fa1, double &fb0, double &fb1) {
#pragma clang fp contract(off)
  if (!have_a) eval64_pair(a, v, fa0, fa1);
  eval64_pair(b, v, fb0, fb1);
  const double ta0 = (fa0 + fb0) / 2.0, ta1 = (fa1 + fb1) / 2.0;
  double pas = b - a;
  double s0 = ta0, s1 = ta1;  // ta + sum of interior points so far
  double R0[ROMB_ITERMAX + 2], R1[ROMB_ITERMAX + 2];
  R0[0] = ta0 * pas;
  R1[0] = ta1 * pas;
  double res0 = R0[0], res1 = R1[0];
  bool run0 = true, run1 = true;
#pragma unroll
  for (int n = 1; n <= ROMB_ITERMAX + 1; n++) {
    if (!(run0 || run1)) break;
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
    double prv0 = R0[0], prv1 = R1[0], top0 = 0.0, top1 = 0.0;
    R0[0] = s0 * pas;
    R1[0] = s1 * pas;
    double r = 1.0;
#pragma unroll
    for (int ii = 1; ii <= ROMB_ITERMAX + 1; ii++) {
      if (ii > n) break;
      r *= 4.0;
      const double o0 = R0[ii], o1 = R1[ii];
      if (ii == n) {
        top0 = prv0;  // t[n-1][0] of the previous level
        top1 = prv1;
      }
      R0[ii] = (r * R0[ii - 1] - prv0) / (r - 1.0);
      R1[ii] = (r * R1[ii - 1] - prv1) / (r - 1.0);
      prv0 = o0;
      prv1 = o1;
    }
    if (run0) {
      res0 = R0[n];
      const double ob = fabs(R0[n] - top0) / R0[n];
      run0 = (ob > ROMB_PREC) && (n <= ROMB_ITERMAX);
    }
    if (run1) {
      res1 = R1[n];
      const double ob = fabs(R1[n] - top1) / R1[n];
      run1 = (ob > ROMB_PREC) && (n <= ROMB_ITERMAX);
    }
  }
  return res0 + res1;
}

// ---------------------------------------------------------------------------
// FP32 integrand in bin-local coordinates
// ---------------------------------------------------------------------------

RX_HD inline float rx_rsqrtf(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / sqrtf(x);
#endif
}

// Interval [rlo, rhi] of one bin integral: g*(s) = ga + s dg from the low end,
// gb - (1-s) dg from the high end (whichever is closer keeps the relative error
// of g* and 1-g* at FP32 rounding), E(s) = rlo + s dE.
struct BinF {
  double rlo, rhi;  // FP64 ends (fallback path)
  float ga, gb, ua, ub, dg, e0, eb, de;
};

// ga, gb: the ends' g* = (E - gmin) del_g, computed by the caller
RX_HD inline BinF make_binf(double rlo, double rhi, double ga, double gb) {
  BinF b;
  b.rlo = rlo;
  b.rhi = rhi;
  b.ga = (float)ga;
  b.gb = (float)gb;
  b.ua = (float)(1.0 - ga);
  b.ub = (float)(1.0 - gb);
  b.dg = (float)(gb - ga);
  b.e0 = (float)rlo;
  b.eb = (float)rhi;
  b.de = (float)(rhi - rlo);
  return b;
}

// Bracket, interpolation factor and E^3 / sqrt(g* (1 - g*)) at s (exact dyadic
// in [0,1]); coordinates from the nearer end, so s = 1 gives exactly the
// values the next bin sees at its s = 0.  *near: the FP32 g* lies within a
// few rounding errors of a g* node, where eval32_fix decides the bracket.
RX_HD inline void eval32_point(float s, const BinF &bf, const RadiusView &v, int &ind, float &inte, float &base,
                               bool &near) {
  float t, u, e;
  if (s <= 0.5f) {
    t = fmaf(s, bf.dg, bf.ga);
    u = fmaf(-s, bf.dg, bf.ua);
    e = fmaf(s, bf.de, bf.e0);
  } else {
    const float sr = 1.0f - s;
    t = fmaf(-sr, bf.dg, bf.gb);
    u = fmaf(sr, bf.dg, bf.ub);
    e = fmaf(-sr, bf.de, bf.eb);
  }
  const float inv_step = (float)v.inv_step;
  const float q = (t - (float)GFAC_H) * inv_step;
  const float thr = 1e-6f * (inv_step + 1.0f);
  ind = min(max((int)floorf(q), 0), v.ng - 2);
  inte = q - (float)ind;
  near = fabsf(q - rintf(q)) < thr;
  base = e * e * e * rx_rsqrtf(t * u);
}

// A point close to a g* node: bracket, factor and weight from the FP64
// coordinates, so that the bracket equals the FP64 path's.
RX_HD inline void eval32_fix(float s, const BinF &bf, const RadiusView &v, int &ind, float &inte, float &base) {
  const double eg = bf.rlo + (bf.rhi - bf.rlo) * (double)s;
  const double egs = (eg - v.gmin) * v.del_g;
  ind = g_bracket(egs, v.ng, v.gs_step, v.inv_step);
  inte = (float)((egs - gstar_node(v.gs_step, ind)) * v.inv_step);
  const float t = (float)egs, u = (float)(1.0 - egs), e = (float)eg;
  base = e * e * e * rx_rsqrtf(t * u);
}

// Both branches at K points: all brackets first (one branch for the rare
// points next to a g* node), then the row reads and the interpolation in one
// straight run (the LDS reads of the K points overlap).
template <int K>
RX_HD inline void eval32_batch(const float (&s)[K], const BinF &bf, const RadiusView &v, float (&f0)[K],
                               float (&f1)[K]) {
  int ind[K];
  float inte[K], base[K];
  bool near[K], any = false;
#pragma unroll
  for (int k = 0; k < K; k++) {
    eval32_point(s[k], bf, v, ind[k], inte[k], base[k], near[k]);
    any = any || near[k];
  }
  if (any) {
#pragma unroll
    for (int k = 0; k < K; k++)
      if (near[k]) eval32_fix(s[k], bf, v, ind[k], inte[k], base[k]);
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    const float a0 = v.T32[ind[k] * 2 + 0], a1 = v.T32[ind[k] * 2 + 1];
    const float b0 = v.T32[ind[k] * 2 + 2], b1 = v.T32[ind[k] * 2 + 3];
    f0[k] = base[k] * fmaf(inte[k], a0 - b0, b0);
    f1[k] = base[k] * fmaf(inte[k], a1 - b1, b1);
  }
  if (v.limb) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      const float c0 = v.C32[ind[k] * 2 + 0], c1 = v.C32[ind[k] * 2 + 1];
      const float d0 = v.C32[ind[k] * 2 + 2], d1 = v.C32[ind[k] * 2 + 3];
      f0[k] *= limb_factor32(v.limb, fmaf(inte[k], c0 - d0, d0));
      f1[k] *= limb_factor32(v.limb, fmaf(inte[k], c1 - d1, d1));
    }
  }
}

// Both branches at one point.
RX_HD inline void eval32s(float s, const BinF &bf, const RadiusView &v, float &f0, float &f1) {
  const float sk[1] = {s};
  float g0[1], g1[1];
  eval32_batch<1>(sk, bf, v, g0, g1);
  f0 = g0[0];
  f1 = g1[0];
}

// Stop test of the reference, ob = |R - top| / R > prec (continue), written
// without the division: ob > prec  <=>  (R > 0 and |d| > prec R) or (R == 0
// and d != 0, ob = inf); R < 0 or NaN gives ob <= 0 or NaN (stop).  A test
// inside the guard band |ob - prec| <= prec guard (R > 0) sets *need64.
RX_HD inline bool romb_continue32(float R, float top, float band, int n, bool *need64) {
  const float d = fabsf(R - top);
  if (n > ROMB_ITERMAX) return false;
  if (R > 0.0f) {
    const float x = fmaf(-(float)ROMB_PREC, R, d);  // d - prec R, one rounding
    if (fabsf(x) <= band * R) *need64 = true;
    return x > 0.0f;
  }
  return R == 0.0f && d > 0.0f;
}

// Romberg for both branches, FP32 values.  Sets *need64 when a stop decision
// fell inside the guard band; the caller then recomputes in FP64.  With
// have_a, (fa0, fa1) is the pair at s = 0; the pair at s = 1 is returned in
// (fb0, fb1).
RX_HD inline float romberg32_pair(const BinF &bf, const RadiusView &v, bool *need64, bool have_a, float fa0,
                                  float fa1, float &fb0, float &fb1) {
  float fm0, fm1;
  // the points of levels 0 and 1 are always needed: evaluate them together
  if (!have_a) eval32s(0.0f, bf, v, fa0, fa1);
  {
    const float sk[2] = {1.0f, 0.5f};
    float g0[2], g1[2];
    eval32_batch<2>(sk, bf, v, g0, g1);
    fb0 = g0[0];
    fb1 = g1[0];
    fm0 = g0[1];
    fm1 = g1[1];
  }
  const float ta0 = (fa0 + fb0) * 0.5f, ta1 = (fa1 + fb1) * 0.5f;
  float pas = bf.de;
  float s0 = ta0, s1 = ta1;
  float R0[ROMB_ITERMAX + 2], R1[ROMB_ITERMAX + 2];
  R0[0] = ta0 * pas;
  R1[0] = ta1 * pas;
  float res0 = R0[0], res1 = R1[0];
  bool run0 = true, run1 = true;
  const float band = (float)ROMB_PREC * v.guard;  // the guard band |ob - prec| <= prec guard
#pragma unroll
  for (int n = 1; n <= ROMB_ITERMAX + 1; n++) {
    if (!(run0 || run1)) break;
    pas = pas * 0.5f;
    const int nnew = 1 << (n - 1);
    const float ds = 1.0f / (float)(2 * nnew);
    float add0 = 0.0f, add1 = 0.0f;
    if (n == 1) {
      add0 += fm0;
      add1 += fm1;
    } else if (n == 2) {
      const float sk[2] = {0.25f, 0.75f};
      float g0[2], g1[2];
      eval32_batch<2>(sk, bf, v, g0, g1);
      add0 = (add0 + g0[0]) + g0[1];
      add1 = (add1 + g1[0]) + g1[1];
    } else {
      // (2m+1) ds, exact: dyadic with at most 7 bits; 4 points per batch
#pragma unroll 1
      for (int m0 = 0; m0 < nnew; m0 += 4) {
        float sk[4], g0[4], g1[4];
#pragma unroll
        for (int k = 0; k < 4; k++) sk[k] = (float)(2 * (m0 + k) + 1) * ds;
        eval32_batch<4>(sk, bf, v, g0, g1);
#pragma unroll
        for (int k = 0; k < 4; k++) {
          add0 += g0[k];
          add1 += g1[k];
        }
      }
    }
    s0 += add0;
    s1 += add1;
    float prv0 = R0[0], prv1 = R1[0], top0 = 0.0f, top1 = 0.0f;
    R0[0] = s0 * pas;
    R1[0] = s1 * pas;
    float r = 1.0f;
#pragma unroll
    for (int ii = 1; ii <= ROMB_ITERMAX + 1; ii++) {
      if (ii > n) break;
      r *= 4.0f;
      const float inv = 1.0f / (r - 1.0f);
      const float o0 = R0[ii], o1 = R1[ii];
      if (ii == n) {
        top0 = prv0;
        top1 = prv1;
      }
      R0[ii] = (r * R0[ii - 1] - prv0) * inv;
      R1[ii] = (r * R1[ii - 1] - prv1) * inv;
      prv0 = o0;
      prv1 = o1;
    }
    if (run0) {
      res0 = R0[n];
      run0 = romb_continue32(R0[n], top0, band, n, need64);
    }
    if (run1) {
      res1 = R1[n];
      run1 = romb_continue32(R1[n], top1, band, n, need64);
    }
  }
  return res0 + res1;
}

// Integral of one energy bin [rlo0, rhi0] (line-energy units) at one radius,
// in reduced units (multiply by emis/(gmax-gmin) for relb_func units).
// glo_u, ghi_u: the ends' unclamped g* = (E - gmin) del_g (a caller walking
// the bins of one radius carries the shared edge from bin to bin).
// MODE 0 (FP64) only; the FP32 path is integ_bin32.
// cy (optional): the integrand pair at rlo0 when cy->ok; on return cy holds the
// pair at rhi0 when this bin evaluated it there (a Romberg bin whose upper end
// is the bin edge), else cy->ok is false.
template <int MODE, class TV = double>
RX_HD inline double integ_bin_g(double rlo0, double rhi0, double glo_u, double ghi_u, const RadiusView &v,
                                bool *need64 = nullptr, EdgePair<TV> *cy = nullptr) {
#pragma clang fp contract(off)
  static_assert(MODE == 0, "the FP32 bin integral is integ_bin32");
  (void)need64;
  const bool have_a = cy && cy->ok;
  if (cy) cy->ok = false;
  // clamp to [0, 1] (the ends are finite)
  const double gblo = fmin(fmax(glo_u, 0.0), 1.0);
  const double gbhi = fmin(fmax(ghi_u, 0.0), 1.0);
  if (gbhi == 0) return 0.0;
  double rlo = rlo0, rhi = rhi0;
  double ga = glo_u, gb = ghi_u;
  double flu = 0.0;
  const bool lo_edge = (gblo <= GFAC_H), hi_edge = (gbhi >= (1.0 - GFAC_H));
  if (gblo <= GFAC_H) {
    double hlo = gblo, hhi = GFAC_H;
    rlo = (GFAC_H * (v.gmax - v.gmin) + v.gmin) * 1.0;
    ga = (rlo - v.gmin) * v.del_g;
    if (gbhi <= GFAC_H) {
      hhi = gbhi;
      rlo = -1.0;
    }
    flu = flu + int_edge_pre(hlo, hhi, v);
  }
  if (gbhi >= (1.0 - GFAC_H)) {
    double hhi = gbhi, hlo = 1.0 - GFAC_H;
    rhi = ((1 - GFAC_H) * (v.gmax - v.gmin) + v.gmin) * 1.0;
    gb = (rhi - v.gmin) * v.del_g;
    if (gblo >= (1.0 - GFAC_H)) {
      hlo = gblo;
      rhi = -1.0;
    }
    flu = flu + int_edge_pre(hlo, hhi, v);
  }
  if ((rhi >= 0) && (rlo >= 0)) {
    if (rlo >= 1.0 * RED_WING) {
      const bool use_a = have_a && !lo_edge;  // the lower end is the shared bin edge
      TV fa0 = use_a ? cy->f0 : (TV)0, fa1 = use_a ? cy->f1 : (TV)0, fb0, fb1;
      flu = flu + romberg64_pair(rlo, rhi, v, use_a, fa0, fa1, fb0, fb1);
      if (cy && !hi_edge) {
        cy->f0 = fb0;
        cy->f1 = fb1;
        cy->ok = true;
      }
    } else {
      const double mid = (rhi + rlo) / 2.0;
      double f0, f1;
      eval64_pair(mid, v, f0, f1);
      flu = flu + (f0 * (rhi - rlo) + f1 * (rhi - rlo));
    }
  }
  return flu;
}

// int_edge for the FP32 path: the bounds and their difference stay FP64 (no
// cancellation: sqrt(hi) - sqrt(lo) = (hi - lo) / (sqrt(hi) + sqrt(lo))), the
// square roots, the quotient and the per-radius factor v.eL / v.eU are FP32.
RX_HD inline float int_edge32(double blo, double bhi, const RadiusView &v) {
  double lo, hi;
  float k;
  if (blo <= 0.5) {
    lo = blo;
    hi = bhi;
    k = v.eL;
  } else {
    lo = 1.0 - bhi;
    hi = 1.0 - blo;
    k = v.eU;
  }
  const float d = (float)(hi - lo);
#if defined(__HIP_DEVICE_COMPILE__)
  const float sq = __builtin_amdgcn_sqrtf((float)hi) + __builtin_amdgcn_sqrtf((float)lo);
#else
  const float sq = sqrtf((float)hi) + sqrtf((float)lo);
#endif
  return (sq > 0.0f) ? k * (d * rx_rcpf(sq)) : 0.0f;
}

// The FP32 path of integ_bin_g (MODE 1) with the value kept in FP32: every
// decision (clamps, edge tests, red-wing switch) is the same FP64 test, the
// edge parts come from int_edge32 and the sum of the parts is FP32.
RX_HD inline float integ_bin32(double rlo0, double rhi0, double glo_u, double ghi_u, const RadiusView &v,
                               bool *need64, EdgePair<float> *cy) {
#pragma clang fp contract(off)
  const bool have_a = cy->ok;
  cy->ok = false;
  const double gblo = fmin(fmax(glo_u, 0.0), 1.0);
  const double gbhi = fmin(fmax(ghi_u, 0.0), 1.0);
  if (gbhi == 0) return 0.0f;
  double rlo = rlo0, rhi = rhi0;
  double ga = glo_u, gb = ghi_u;
  float flu = 0.0f;
  const bool lo_edge = (gblo <= GFAC_H), hi_edge = (gbhi >= (1.0 - GFAC_H));
  if (lo_edge) {
    double hlo = gblo, hhi = GFAC_H;
    rlo = (GFAC_H * (v.gmax - v.gmin) + v.gmin) * 1.0;
    ga = (rlo - v.gmin) * v.del_g;
    if (gbhi <= GFAC_H) {
      hhi = gbhi;
      rlo = -1.0;
    }
    flu += int_edge32(hlo, hhi, v);
  }
  if (hi_edge) {
    double hhi = gbhi, hlo = 1.0 - GFAC_H;
    rhi = ((1 - GFAC_H) * (v.gmax - v.gmin) + v.gmin) * 1.0;
    gb = (rhi - v.gmin) * v.del_g;
    if (gblo >= (1.0 - GFAC_H)) {
      hlo = gblo;
      rhi = -1.0;
    }
    flu += int_edge32(hlo, hhi, v);
  }
  if ((rhi >= 0) && (rlo >= 0)) {
    const BinF bf = make_binf(rlo, rhi, ga, gb);
    if (rlo >= 1.0 * RED_WING) {
      const bool use_a = have_a && !lo_edge;  // the lower end is the shared bin edge
      float fb0, fb1;
      flu += romberg32_pair(bf, v, need64, use_a, use_a ? cy->f0 : 0.0f, use_a ? cy->f1 : 0.0f, fb0, fb1);
      if (!hi_edge) {
        cy->f0 = fb0;
        cy->f1 = fb1;
        cy->ok = true;
      }
    } else {
      float f0, f1;
      eval32s(0.5f, bf, v, f0, f1);
      flu += (f0 + f1) * bf.de;
    }
  }
  return flu;
}

template <int MODE>
RX_HD inline double integ_bin(double rlo0, double rhi0, const RadiusView &v, bool *need64 = nullptr) {
#pragma clang fp contract(off)
  return integ_bin_g<MODE>(rlo0, rhi0, (rlo0 / 1.0 - v.gmin) * v.del_g, (rhi0 / 1.0 - v.gmin) * v.del_g, v,
                           need64);
}

// ---- small helpers shared by setup code ---------------------------------

RX_HD inline double lerp1(double ifac, double lo, double hi) { return lerp1_(ifac, lo, hi); }

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

// fine radial grid point ii (src/relutility.c:666-677), from r1 = 1/sqrt(Rout)
// and r2 = 1/sqrt(Rin)
RX_HD inline double fine_radius_at(double r1, double r2, int ii, int nr) {
#pragma clang fp contract(off)
  double x = ((double)(ii)) * (r2 - r1) / (nr - 1) + r1;
  const double y = 1.0 / x;
  return y * y;
}
RX_HD inline double fine_radius_r1(double rout) { return 1.0 / sqrt(rout); }
RX_HD inline double fine_radius(double rin, double rout, int ii, int nr) {
  return fine_radius_at(fine_radius_r1(rout), fine_radius_r1(rin), ii, nr);
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
