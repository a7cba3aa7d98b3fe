// tables.cpp -- synthetic relativistic tables and the RXTB file format.
//
// The fields follow the Newtonian + Doppler thin-disk picture:
//   g0 = 1/u^t(r, a) for a Keplerian Kerr disk (src/Relphysics.cpp:28-31),
//   orbital speed v(r, a) in the ZAMO frame, projected as beta = v sin(i),
//   gmin = g0/(1+beta), gmax = g0/(1-beta);
//   trff ~ (2/pi) cos(i) (1 + 0.6/r), so that an outer ring's line flux tends
//   to 0.5 cos(i) of its area fraction (the convergence the reference's
//   normalisation test checks, test/unit/test-stdfunctions.cpp:214-251);
//   cosne ~ cos(i) bent towards 1 at small radii;
//   lamp-post intensity = Newtonian point-source emissivity
//   (src/Relphysics.cpp:282-288) times a smooth GR-like boost, emission and
//   incidence angles ~ atan(r/h).
// Each field carries a seeded smooth perturbation (a few low-frequency
// sinusoids), so the tables are deterministic functions of (seed, dims).
#include "rx_tables.h"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace rx {

namespace {

constexpr double kPi = 3.14159265358979323846;

struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ULL + 0x1234567ULL) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Sum of four sinusoids in up to four normalised coordinates, |f| <= 1.
struct Smooth {
  double amp[4], fr[4][4], ph[4];
  explicit Smooth(SplitMix &rng) {
    double tot = 0;
    for (int m = 0; m < 4; m++) {
      amp[m] = 0.2 + rng.uniform();
      tot += amp[m];
      for (int d = 0; d < 4; d++) fr[m][d] = 0.3 + 1.2 * rng.uniform();
      ph[m] = 2 * kPi * rng.uniform();
    }
    for (int m = 0; m < 4; m++) amp[m] /= tot;
  }
  double operator()(double x0, double x1 = 0, double x2 = 0, double x3 = 0) const {
    double v = 0;
    for (int m = 0; m < 4; m++) {
      v += amp[m] * std::sin(2 * kPi * (fr[m][0] * x0 + fr[m][1] * x1 + fr[m][2] * x2 +
                                        fr[m][3] * x3) + ph[m]);
    }
    return v;
  }
};

// spin grid on [-0.998, 0.998], denser towards +1
std::vector<float> spin_grid(int n) {
  std::vector<float> a(n);
  for (int i = 0; i < n; i++) {
    double t = (n > 1) ? (double)i / (n - 1) : 0.0;
    a[i] = (float)(0.998 * (1.0 - 2.0 * std::pow(1.0 - t, 1.6)));
  }
  a[0] = -0.998f;
  a[n - 1] = 0.998f;
  return a;
}

// orbital speed of a circular equatorial orbit seen by a ZAMO (Bardeen+72)
double orbit_speed(double r, double a) {
  double sr = std::sqrt(r);
  double delta = r * r - 2 * r + a * a;
  return (r * r - 2 * a * sr + a * a) / (std::sqrt(delta) * (r * sr + a));
}

double ut_disk(double r, double a) {
  return ((r * std::sqrt(r) + a) / (std::sqrt(r) * std::sqrt(r * r - 3 * r + 2 * a * std::sqrt(r))));
}

}  // namespace

double kerr_rms(double a) {
  double sgn = (a < 0) ? -1.0 : 1.0;
  double z1 = 1.0 + std::pow(1.0 - a * a, 1.0 / 3.0) *
                        (std::pow(1.0 + a, 1.0 / 3.0) + std::pow(1.0 - a, 1.0 / 3.0));
  double z2 = std::sqrt(3.0 * a * a + z1 * z1);
  return 3.0 + z2 - sgn * std::sqrt((3.0 - z1) * (3.0 + z1 + 2 * z2));
}

double kerr_rplus(double a) { return 1 + std::sqrt(1 - a * a); }

RelTableHost make_synthetic_reltable(uint64_t seed, int n_a, int n_mu0, int n_r, int n_g) {
  RelTableHost t;
  t.n_a = n_a;
  t.n_mu0 = n_mu0;
  t.n_r = n_r;
  t.n_g = n_g;
  SplitMix rng(seed);
  Smooth s_gmin(rng), s_gmax(rng), s_t1(rng), s_t2(rng), s_c1(rng), s_c2(rng);

  t.a = spin_grid(n_a);
  t.mu0.resize(n_mu0);
  // inclinations from 87.1 deg down to 2.9 deg: covers the allowed 3-87 deg
  for (int j = 0; j < n_mu0; j++) {
    double th = (87.1 - (87.1 - 2.9) * j / (n_mu0 - 1)) * kPi / 180.0;
    t.mu0[j] = (float)std::cos(th);
  }
  const size_t nn = t.node_count();
  t.r.resize(nn * n_r);
  t.gmin.resize(nn * n_r);
  t.gmax.resize(nn * n_r);
  t.trff1.resize(nn * n_r * n_g);
  t.trff2.resize(nn * n_r * n_g);
  t.cosne1.resize(nn * n_r * n_g);
  t.cosne2.resize(nn * n_r * n_g);

  const double rmax = 999.0;  // within 1% of 1000 rg, as the reference expects
  for (int ia = 0; ia < n_a; ia++) {
    const double a = t.a[ia];
    const double rmin = kerr_rms(a);
    const double ta = (double)ia / (n_a - 1);
    for (int im = 0; im < n_mu0; im++) {
      const double mu0 = t.mu0[im];
      const double sini = std::sqrt(1.0 - mu0 * mu0);
      const double tm = (double)im / (n_mu0 - 1);
      const size_t node = (size_t)ia * n_mu0 + im;
      for (int ir = 0; ir < n_r; ir++) {
        const double u = (double)ir / (n_r - 1);
        const double x = u * (1.0 / std::sqrt(rmin) - 1.0 / std::sqrt(rmax)) + 1.0 / std::sqrt(rmax);
        double r = 1.0 / (x * x);
        if (ir == n_r - 1) r = rmin;
        const float rf = (float)r;
        const double rr = (double)rf;
        t.r[node * n_r + ir] = rf;
        const double g0 = 1.0 / ut_disk(std::max(rr, rmin), a);
        const double beta = orbit_speed(std::max(rr, rmin), a) * sini;
        double gmn = g0 / (1.0 + beta) * (1.0 + 0.01 * s_gmin(ta, tm, u));
        double gmx = g0 / (1.0 - beta) * (1.0 + 0.01 * s_gmax(ta, tm, u));
        if (gmx < gmn * 1.0005) gmx = gmn * 1.0005;
        t.gmin[node * n_r + ir] = (float)gmn;
        t.gmax[node * n_r + ir] = (float)gmx;
        const double base = (2.0 / kPi) * mu0 * (1.0 + 0.6 / rr);
        const double bend = 0.25 / (1.0 + rr / 6.0);
        for (int ig = 0; ig < n_g; ig++) {
          const double gs = (double)ig / (n_g - 1);
          const size_t o = (node * n_r + ir) * n_g + ig;
          t.trff1[o] = (float)(base * (1.0 - 0.2 * (gs - 0.5)) * (1.0 + 0.03 * s_t1(ta, tm, u, gs)));
          t.trff2[o] = (float)(base * (1.0 + 0.2 * (gs - 0.5)) * (1.0 + 0.03 * s_t2(ta, tm, u, gs)));
          double c1 = mu0 + (1.0 - mu0) * bend * (0.6 + 0.4 * std::sin(kPi * gs)) + 0.02 * s_c1(ta, tm, u, gs);
          double c2 = mu0 + (1.0 - mu0) * bend * (0.6 + 0.4 * std::sin(kPi * gs + 0.7)) + 0.02 * s_c2(ta, tm, u, gs);
          t.cosne1[o] = (float)std::min(1.0, std::max(0.01, c1));
          t.cosne2[o] = (float)std::min(1.0, std::max(0.01, c2));
        }
      }
    }
  }
  return t;
}

LpTableHost make_synthetic_lptable(uint64_t seed, int n_a, int n_h, int n_r) {
  LpTableHost t;
  t.n_a = n_a;
  t.n_h = n_h;
  t.n_r = n_r;
  SplitMix rng(seed ^ 0xA5A5A5A5ULL);
  Smooth s_int(rng), s_del(rng), s_inc(rng);
  t.a = spin_grid(n_a);
  t.h.resize((size_t)n_a * n_h);
  t.rad.resize((size_t)n_a * n_r);
  t.intens.resize((size_t)n_a * n_h * n_r);
  t.del.resize((size_t)n_a * n_h * n_r);
  t.del_inc.resize((size_t)n_a * n_h * n_r);
  for (int ia = 0; ia < n_a; ia++) {
    const double a = t.a[ia];
    const double rp = kerr_rplus(a);
    const double ta = (double)ia / (n_a - 1);
    // per-spin height grid from just below 1.1 r+ to 500 rg (log spaced)
    const double hlo = 1.05 * rp, hhi = 500.0;
    for (int ih = 0; ih < n_h; ih++) {
      t.h[(size_t)ia * n_h + ih] =
          (float)std::exp(std::log(hlo) + (std::log(hhi) - std::log(hlo)) * ih / (n_h - 1));
    }
    // radii from the horizon to exactly 1000 rg (log spaced)
    for (int ir = 0; ir < n_r; ir++) {
      double r = std::exp(std::log(rp) + (std::log(1000.0) - std::log(rp)) * ir / (n_r - 1));
      if (ir == n_r - 1) r = 1000.0;
      t.rad[(size_t)ia * n_r + ir] = (float)r;
    }
    for (int ih = 0; ih < n_h; ih++) {
      const double h = t.h[(size_t)ia * n_h + ih];
      const double th = (double)ih / (n_h - 1);
      for (int ir = 0; ir < n_r; ir++) {
        const double r = t.rad[(size_t)ia * n_r + ir];
        const double tr = (double)ir / (n_r - 1);
        const size_t o = ((size_t)ia * n_h + ih) * n_r + ir;
        const double newton = std::pow(1. / (std::pow(r / h, 2) + 1), 1.5) / (2 * kPi * h * h);
        const double boost = (1.0 + 2.0 / r) * (1.0 + 0.05 * s_int(ta, th, tr));
        t.intens[o] = (float)(newton * boost);
        const double ang = std::atan(r / h);
        const double bend = 0.08 / (1.0 + r / h);
        t.del[o] = (float)std::fabs(ang * (1.0 + bend * (1.0 + 0.2 * s_del(ta, th, tr))));
        t.del_inc[o] = (float)std::fabs(ang * (1.0 + 0.05 * s_inc(ta, th, tr)));
      }
    }
  }
  return t;
}

std::string validate_reltable(const RelTableHost &t) {
  char buf[256];
  if (t.n_a < 2 || t.n_mu0 < 2 || t.n_r < 3 || t.n_g < 2) return "table dims too small";
  for (int i = 0; i + 1 < t.n_a; i++)
    if (!(t.a[i] < t.a[i + 1])) return "spin grid not ascending";
  for (int i = 0; i + 1 < t.n_mu0; i++)
    if (!(t.mu0[i] < t.mu0[i + 1])) return "mu0 grid not ascending";
  if (!(t.mu0[0] > 0.0f) || !(t.mu0[t.n_mu0 - 1] <= 1.0f)) return "mu0 grid outside (0,1]";
  for (size_t n = 0; n < t.node_count(); n++) {
    for (int i = 0; i < t.n_r; i++) {
      const size_t o = n * t.n_r + i;
      if (i + 1 < t.n_r && !(t.r[o] > t.r[o + 1])) {
        snprintf(buf, sizeof buf, "radii not descending at node %zu index %d", n, i);
        return buf;
      }
      if (!(t.gmin[o] > 0.0f) || !(t.gmax[o] > t.gmin[o])) return "gmin/gmax invalid";
      for (int g = 0; g < t.n_g; g++) {
        const size_t q = o * t.n_g + g;
        if (!(t.trff1[q] >= 0.0f) || !(t.trff2[q] >= 0.0f)) return "trff negative";
        if (!(t.cosne1[q] > 0.0f && t.cosne1[q] <= 1.0f) || !(t.cosne2[q] > 0.0f && t.cosne2[q] <= 1.0f))
          return "cosne outside (0,1]";
      }
    }
    if (!(t.r[n * t.n_r] <= 1000.0f)) return "outer radius above 1000 rg";
  }
  return "";
}

std::string validate_lptable(const LpTableHost &t) {
  if (t.n_a < 2 || t.n_h < 2 || t.n_r < 2) return "lp table dims too small";
  for (int i = 0; i + 1 < t.n_a; i++)
    if (!(t.a[i] < t.a[i + 1])) return "lp spin grid not ascending";
  for (int ia = 0; ia < t.n_a; ia++) {
    for (int i = 0; i + 1 < t.n_h; i++)
      if (!(t.h[(size_t)ia * t.n_h + i] < t.h[(size_t)ia * t.n_h + i + 1])) return "lp h grid not ascending";
    for (int i = 0; i + 1 < t.n_r; i++)
      if (!(t.rad[(size_t)ia * t.n_r + i] < t.rad[(size_t)ia * t.n_r + i + 1])) return "lp radii not ascending";
    for (size_t q = (size_t)ia * t.n_h * t.n_r; q < (size_t)(ia + 1) * t.n_h * t.n_r; q++) {
      if (!(t.intens[q] > 0.0f)) return "lp intensity not positive";
      if (!(t.del[q] >= 0.0f) || !(t.del_inc[q] >= 0.0f)) return "lp angles negative";
    }
  }
  return "";
}

// ---- xillver reflection table ----

namespace {
// values of one synthetic parameter axis: n points on [lo, hi], linear or log
std::vector<float> axis(int n, double lo, double hi, bool log_spaced) {
  std::vector<float> v(n);
  for (int i = 0; i < n; i++) {
    const double t = (n > 1) ? (double)i / (n - 1) : 0.0;
    v[i] = (float)(log_spaced ? lo * std::pow(hi / lo, t) : lo + (hi - lo) * t);
  }
  return v;
}
}  // namespace

// Synthetic xillver spectra (photons per bin): a cut-off power law
// E^(1-Gamma) exp(-E/Ecut) dE with a seeded smooth soft-band modulation that
// changes with logXi, a limb-brightening factor in the emission angle and a
// Gaussian Fe K line at 6.4 keV whose strength scales with A_Fe; every row is
// multiplied by 10^logXi (the scale in which xillver tables store spectra).
// Inclinations are the centres of n_incl equal bins in cos(i), as the angle
// bins of the relativistic profile (src/Relprofile.cpp:799-801).
XillTableHost make_synthetic_xilltable(uint64_t seed, int n_gam, int n_afe, int n_lxi, int n_ect,
                                       int n_incl, int n_ener, double e4_lo, double e4_hi) {
  XillTableHost t;
  t.n_par = 5;
  const int ids[5] = {XP_GAM, XP_AFE, XP_LXI, XP_ECT, XP_INC};
  const int nv[5] = {n_gam, n_afe, n_lxi, n_ect, n_incl};
  for (int k = 0; k < 5; k++) {
    t.pid[k] = ids[k];
    t.n_val[k] = nv[k];
  }
  t.val[0] = axis(n_gam, 1.0, 3.4, false);
  t.val[1] = axis(n_afe, 0.5, 10.0, true);
  t.val[2] = axis(n_lxi, 0.0, 4.7, false);
  t.val[3] = axis(n_ect, e4_lo, e4_hi, true);
  t.val[4].resize(n_incl);
  for (int i = 0; i < n_incl; i++) t.val[4][i] = (float)(std::acos(1.0 - (i + 0.5) / n_incl) * 180.0 / kPi);
  t.n_ener = n_ener;
  t.elo.resize(n_ener);
  t.ehi.resize(n_ener);
  std::vector<double> lne(n_ener), emid(n_ener), de(n_ener), line(n_ener), soft(n_ener);
  SplitMix rng(seed);
  Smooth s_soft(rng), s_pl(rng);
  for (int e = 0; e < n_ener; e++) {
    const double lo = 0.07 * std::pow(1000.0 / 0.07, (double)e / n_ener);
    const double hi = 0.07 * std::pow(1000.0 / 0.07, (double)(e + 1) / n_ener);
    t.elo[e] = (float)lo;
    t.ehi[e] = (float)hi;
    emid[e] = 0.5 * (lo + hi);
    de[e] = hi - lo;
    lne[e] = std::log(emid[e]);
    const double x = (lne[e] - std::log(6.4)) / 0.03;  // sigma 3% of 6.4 keV
    line[e] = 3.0 * std::exp(-0.5 * x * x);
    soft[e] = std::exp(-emid[e] / 2.0);
  }
  t.spec.resize(t.n_rows() * (size_t)n_ener);
  std::vector<double> base(n_ener);
  size_t row = 0;
  for (int ig = 0; ig < n_gam; ig++) {
    const double gam = t.val[0][ig];
    for (int ia = 0; ia < n_afe; ia++) {
      const double afe = t.val[1][ia];
      for (int ix = 0; ix < n_lxi; ix++) {
        const double lxi = t.val[2][ix];
        const double scale = std::pow(10.0, lxi);
        for (int ic = 0; ic < n_ect; ic++) {
          const double ect = t.val[3][ic];
          for (int e = 0; e < n_ener; e++) {
            const double u = lne[e] / std::log(1000.0);
            base[e] = (1.0 - gam) * lne[e] - emid[e] / ect + 0.08 * s_pl(u, gam / 3.4) +
                      0.6 * (lxi / 4.7) * soft[e] * (1.0 + 0.5 * s_soft(u, lxi / 4.7));
          }
          for (int ii = 0; ii < n_incl; ii++, row++) {
            const double mu = std::cos(t.val[4][ii] * kPi / 180.0);
            const double limb = std::log(0.35 + 0.65 * mu);
            float *dst = &t.spec[row * n_ener];
            for (int e = 0; e < n_ener; e++) {
              const double v = scale * de[e] * std::exp(base[e] + limb * (1.0 + 0.05 * lne[e])) *
                               (1.0 + afe * line[e] * (0.6 + 0.4 * mu));
              dst[e] = (float)v;
            }
          }
        }
      }
    }
  }
  return t;
}

// Synthetic spectra of the blackbody models (photons per bin): the reflected
// blackbody E^2 / (exp(E / kT) - 1) softened by a power-law wing, a smooth
// ionisation and density dependent soft excess, an iron line growing with
// A_Fe and the limb law of the other table; 10^logXi times the density's
// 10^(logN - 15) as the files store them (the loader divides both out).
XillTableHost make_synthetic_xilltable_ns(uint64_t seed, int n_ktbb, int n_afe, int n_lxi, int n_dens, int n_incl,
                                          int n_ener) {
  XillTableHost t;
  t.n_par = 5;
  const int ids[5] = {XP_KTB, XP_AFE, XP_LXI, XP_DNS, XP_INC};
  const int nv[5] = {n_ktbb, n_afe, n_lxi, n_dens, n_incl};
  for (int k = 0; k < 5; k++) {
    t.pid[k] = ids[k];
    t.n_val[k] = nv[k];
  }
  t.val[0] = axis(n_ktbb, 0.5, 10.0, true);
  t.val[1] = axis(n_afe, 0.5, 10.0, true);
  t.val[2] = axis(n_lxi, 1.0, 4.7, false);
  t.val[3] = axis(n_dens, 15.0, 19.0, false);
  t.val[4].resize(n_incl);
  for (int i = 0; i < n_incl; i++) t.val[4][i] = (float)(std::acos(1.0 - (i + 0.5) / n_incl) * 180.0 / kPi);
  t.n_ener = n_ener;
  t.elo.resize(n_ener);
  t.ehi.resize(n_ener);
  std::vector<double> lne(n_ener), emid(n_ener), de(n_ener), line(n_ener);
  SplitMix rng(seed);
  Smooth s_soft(rng), s_bb(rng);
  for (int e = 0; e < n_ener; e++) {
    const double lo = 0.07 * std::pow(1000.0 / 0.07, (double)e / n_ener);
    const double hi = 0.07 * std::pow(1000.0 / 0.07, (double)(e + 1) / n_ener);
    t.elo[e] = (float)lo;
    t.ehi[e] = (float)hi;
    emid[e] = 0.5 * (lo + hi);
    de[e] = hi - lo;
    lne[e] = std::log(emid[e]);
    const double x = (lne[e] - std::log(6.4)) / 0.03;
    line[e] = 3.0 * std::exp(-0.5 * x * x);
  }
  t.spec.resize(t.n_rows() * (size_t)n_ener);
  std::vector<double> base(n_ener);
  size_t row = 0;
  for (int it = 0; it < n_ktbb; it++) {
    const double kt = t.val[0][it];
    for (int ia = 0; ia < n_afe; ia++) {
      const double afe = t.val[1][ia];
      for (int ix = 0; ix < n_lxi; ix++) {
        const double lxi = t.val[2][ix];
        for (int id = 0; id < n_dens; id++) {
          const double dens = t.val[3][id];
          const double scale = std::pow(10.0, lxi) * std::pow(10.0, dens - 15.0);
          for (int e = 0; e < n_ener; e++) {
            const double u = lne[e] / std::log(1000.0);
            const double y = std::min(emid[e] / kt, 700.0);
            const double bb = emid[e] * emid[e] / std::expm1(y) + 1e-3 * std::pow(emid[e], -1.5);
            base[e] = std::log(bb) + 0.08 * s_bb(u, kt / 10.0) +
                      0.5 * (lxi / 4.7) * std::exp(-emid[e]) * (1.0 + 0.5 * s_soft(u, lxi / 4.7, (dens - 15) / 4));
          }
          for (int ii = 0; ii < n_incl; ii++, row++) {
            const double mu = std::cos(t.val[4][ii] * kPi / 180.0);
            const double limb = std::log(0.35 + 0.65 * mu);
            float *dst = &t.spec[row * n_ener];
            for (int e = 0; e < n_ener; e++) {
              const double v = scale * de[e] * std::exp(base[e] + limb * (1.0 + 0.05 * lne[e])) *
                               (1.0 + afe * line[e] * (0.6 + 0.4 * mu));
              dst[e] = (float)v;
            }
          }
        }
      }
    }
  }
  return t;
}

// A six-parameter table: the axes of make_synthetic_xilltable with Dens
// (15-19) between the fourth (Ecut / kTe, which stays the table's axis 3,
// the one ensure_ecut_within_boundarys reads) and the inclination; each
// density's rows are that table's rows times 10^(logN - 15) (the scale the
// files store) and a density-dependent soft excess, so that the density
// changes the spectra.  Gamma is the first axis, interpolated outside the
// other five.
XillTableHost make_synthetic_xilltable_dens(uint64_t seed, int n_dens, int n_gam, int n_afe, int n_lxi, int n_ect,
                                            int n_incl, int n_ener, double e4_lo, double e4_hi) {
  const XillTableHost b = make_synthetic_xilltable(seed, n_gam, n_afe, n_lxi, n_ect, n_incl, n_ener, e4_lo, e4_hi);
  XillTableHost t;
  t.n_par = 6;
  for (int k = 0; k < 4; k++) {
    t.pid[k] = b.pid[k];
    t.n_val[k] = b.n_val[k];
    t.val[k] = b.val[k];
  }
  t.pid[4] = XP_DNS;
  t.n_val[4] = n_dens;
  t.val[4] = axis(n_dens, 15.0, 19.0, false);
  t.pid[5] = b.pid[4];
  t.n_val[5] = b.n_val[4];
  t.val[5] = b.val[4];
  t.n_ener = n_ener;
  t.elo = b.elo;
  t.ehi = b.ehi;
  t.spec.resize(b.spec.size() * (size_t)n_dens);
  std::vector<double> soft(n_ener);
  for (int e = 0; e < n_ener; e++) soft[e] = std::exp(-0.5 * ((double)b.elo[e] + (double)b.ehi[e]) / 1.5);
  const size_t n_outer = (size_t)n_gam * n_afe * n_lxi * n_ect;  // rows of b without the inclination
  for (size_t q = 0; q < n_outer; q++)
    for (int id = 0; id < n_dens; id++) {
      const double x = (t.val[4][id] - 15.0) / 4.0;
      const double scale = std::pow(10.0, t.val[4][id] - 15.0);
      for (int ii = 0; ii < n_incl; ii++) {
        const float *src = &b.spec[(q * n_incl + ii) * (size_t)n_ener];
        float *dst = &t.spec[((q * n_dens + id) * n_incl + ii) * (size_t)n_ener];
        for (int e = 0; e < n_ener; e++) dst[e] = (float)((double)src[e] * scale * (1.0 + 0.4 * x * soft[e]));
      }
    }
  return t;
}

std::string validate_xilltable(const XillTableHost &t) {
  if (t.n_par != 5 && t.n_par != 6) return "xillver table must have five or six parameters";
  if (t.pid[t.n_par - 1] != XP_INC) return "the last xillver parameter must be the inclination";
  // a sixth parameter is interpolated outside the other five; the Ecut / kTe
  // bracket, which the lamp post's zones move, must be an inner one
  if (t.n_par == 6 && t.pid[0] == XP_ECT) return "Ecut / kTe as the first of six xillver parameters is not supported";
  bool seen[8] = {};
  for (int k = 0; k < t.n_par; k++) {
    if (t.pid[k] < 0 || t.pid[k] > XP_INC || seen[t.pid[k]]) return "xillver parameter ids invalid or repeated";
    seen[t.pid[k]] = true;
    if (t.n_val[k] < 2 || (int)t.val[k].size() != t.n_val[k]) return "xillver parameter axis too short";
    for (int i = 0; i + 1 < t.n_val[k]; i++)
      if (!(t.val[k][i] < t.val[k][i + 1])) return "xillver parameter values not ascending";
  }
  if (!seen[XP_LXI]) return "xillver table without a logXi axis";
  if (t.n_ener < 2 || (int)t.elo.size() != t.n_ener || (int)t.ehi.size() != t.n_ener) return "xillver energy grid invalid";
  for (int e = 0; e < t.n_ener; e++) {
    if (!(t.elo[e] > 0.0f) || !(t.ehi[e] > t.elo[e])) return "xillver energy bins invalid";
    if (e + 1 < t.n_ener && !(t.elo[e + 1] >= t.ehi[e])) return "xillver energy bins overlap";
  }
  if (t.spec.size() != t.n_rows() * (size_t)t.n_ener) return "xillver spectra size mismatch";
  for (float v : t.spec)
    if (!(v >= 0.0f) || !std::isfinite(v)) return "xillver spectrum negative or not finite";
  return "";
}

// renorm_xill_spec (src/xilltable.c:478-486): each division is rounded to
// float (the spectrum is a float array); the logXi and density of a row are
// the table values of its indices (normalizeXillverSpecLogxiDensity, :488-511);
// a table without a density axis uses the default 15, which skips that step.
std::vector<float> xilltable_renormed(const XillTableHost &t) {
  std::vector<float> out(t.spec.size());
  const int d_lxi = t.dim_of(XP_LXI), d_dns = t.dim_of(XP_DNS);
  const size_t n_rows = t.n_rows();
  for (size_t row = 0; row < n_rows; row++) {
    size_t rest = row;
    int idx[XILL_NPAR] = {};
    for (int k = t.n_par - 1; k >= 0; k--) {
      idx[k] = (int)(rest % t.n_val[k]);
      rest /= t.n_val[k];
    }
    const double lxi = t.val[d_lxi][idx[d_lxi]];
    const double dens = d_dns >= 0 ? (double)t.val[d_dns][idx[d_dns]] : 15.0;
    const double f_lxi = std::pow(10, lxi);
    const bool do_dens = std::fabs(dens - 15) > 1e-6;
    const double f_dens = do_dens ? std::pow(10, dens - 15) : 1.0;
    const float *src = &t.spec[row * t.n_ener];
    float *dst = &out[row * t.n_ener];
    for (int e = 0; e < t.n_ener; e++) {
      float v = src[e];
      v = (float)(v / f_lxi);
      if (do_dens) v = (float)(v / f_dens);
      dst[e] = v;
    }
  }
  return out;
}

// ---- returning-radiation table ----

// Synthetic returning fractions: n_rad zones log-spaced from the spin's ISCO to
// 1000 rg; the energy shift between two radii centres on the ratio of the disk
// u^t, with a width that grows towards small radii; a smooth, seeded
// distribution over the g bins (sums to 1); the returning fraction falls with
// the distance between the zones and with radius.
RradTableHost make_synthetic_rradtable(uint64_t seed, int n_spin, int n_rad, int n_g) {
  RradTableHost t;
  t.n_spin = n_spin;
  t.n_rad = n_rad;
  t.n_g = n_g;
  t.spin.resize(n_spin);
  for (int k = 0; k < n_spin; k++) t.spin[k] = -0.998 + 1.996 * k / (n_spin - 1);
  const size_t nr2 = (size_t)n_rad * n_rad;
  t.rlo.resize((size_t)n_spin * n_rad);
  t.rhi.resize((size_t)n_spin * n_rad);
  t.f_ret.resize((size_t)n_spin * n_rad);
  t.tf_r.resize(n_spin * nr2);
  t.gmin.resize(n_spin * nr2);
  t.gmax.resize(n_spin * nr2);
  t.frac_g.resize(n_spin * nr2 * n_g);
  SplitMix rng(seed);
  Smooth s_tf(rng), s_g(rng);
  for (int k = 0; k < n_spin; k++) {
    const double a = t.spin[k];
    const double r0 = kerr_rms(a);
    std::vector<double> rm(n_rad);
    for (int i = 0; i < n_rad; i++) {
      t.rlo[(size_t)k * n_rad + i] = r0 * std::pow(1000.0 / r0, (double)i / n_rad);
      t.rhi[(size_t)k * n_rad + i] = r0 * std::pow(1000.0 / r0, (double)(i + 1) / n_rad);
      rm[i] = 0.5 * (t.rlo[(size_t)k * n_rad + i] + t.rhi[(size_t)k * n_rad + i]);
    }
    for (int io = 0; io < n_rad; io++) {
      double fret = 0.0;
      for (int ie = 0; ie < n_rad; ie++) {
        const size_t q = (size_t)k * nr2 + (size_t)io * n_rad + ie;
        const double d = std::log(rm[io] / rm[ie]);
        const double tf = 0.08 * std::exp(-0.8 * std::fabs(d)) * std::pow(rm[ie] / r0, -0.7) *
                          (1.0 + 0.2 * s_tf((double)io / n_rad, (double)ie / n_rad, (a + 1) / 2));
        t.tf_r[q] = tf;
        fret += tf;
        const double g0 = ut_disk(rm[io], a) / ut_disk(rm[ie], a);
        const double w = 0.1 + 0.5 / std::sqrt(rm[ie]);
        t.gmin[q] = g0 * (1.0 - w);
        t.gmax[q] = g0 * (1.0 + w);
        double sum = 0.0;
        for (int g = 0; g < n_g; g++) {
          const double x = (g + 0.5) / n_g;
          const double v = (0.3 + std::sqrt(x * (1.0 - x))) * (1.0 + 0.3 * s_g(x, (double)ie / n_rad, (double)io / n_rad));
          t.frac_g[q * n_g + g] = v;
          sum += v;
        }
        for (int g = 0; g < n_g; g++) t.frac_g[q * n_g + g] /= sum;
      }
      t.f_ret[(size_t)k * n_rad + io] = fret;
    }
  }
  return t;
}

std::string validate_rradtable(const RradTableHost &t) {
  if (t.n_spin < 1 || t.n_rad < 2 || t.n_g < 1) return "returning-radiation table dims too small";
  const size_t nr2 = (size_t)t.n_rad * t.n_rad;
  if (t.spin.size() != (size_t)t.n_spin || t.rlo.size() != (size_t)t.n_spin * t.n_rad ||
      t.tf_r.size() != t.n_spin * nr2 || t.frac_g.size() != t.n_spin * nr2 * t.n_g)
    return "returning-radiation table size mismatch";
  for (int k = 0; k + 1 < t.n_spin; k++)
    if (!(t.spin[k] < t.spin[k + 1])) return "returning-radiation spins not ascending";
  for (int k = 0; k < t.n_spin; k++)
    for (int i = 0; i < t.n_rad; i++) {
      const size_t o = (size_t)k * t.n_rad + i;
      if (!(t.rhi[o] > t.rlo[o]) || (i + 1 < t.n_rad && !(t.rlo[o + 1] >= t.rlo[o])))
        return "returning-radiation radii not ascending";
    }
  for (size_t q = 0; q < t.tf_r.size(); q++)
    if (!(t.tf_r[q] >= 0) || !(t.gmax[q] > t.gmin[q]) || !(t.gmin[q] > 0)) return "returning fractions invalid";
  return "";
}

// ---- RXTB binary format: magic, kind, dims, then the float arrays ----

namespace {
constexpr uint32_t kMagic = 0x42545852u;  // "RXTB"

bool wr(FILE *f, const void *p, size_t n) { return fwrite(p, 1, n, f) == n; }
bool rd(FILE *f, void *p, size_t n) { return fread(p, 1, n, f) == n; }
bool wvec(FILE *f, const std::vector<float> &v) { return wr(f, v.data(), v.size() * sizeof(float)); }
bool rvec(FILE *f, std::vector<float> &v, size_t n) {
  v.resize(n);
  return rd(f, v.data(), n * sizeof(float));
}
}  // namespace

std::string save_reltable(const RelTableHost &t, const std::string &path) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return "cannot open " + path;
  int32_t hdr[6] = {(int32_t)kMagic, 1, t.n_a, t.n_mu0, t.n_r, t.n_g};
  bool ok = wr(f, hdr, sizeof hdr) && wvec(f, t.a) && wvec(f, t.mu0) && wvec(f, t.r) &&
            wvec(f, t.gmin) && wvec(f, t.gmax) && wvec(f, t.trff1) && wvec(f, t.trff2) &&
            wvec(f, t.cosne1) && wvec(f, t.cosne2);
  fclose(f);
  return ok ? "" : "write failed: " + path;
}

std::string load_reltable(RelTableHost &t, const std::string &path) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return "cannot open " + path;
  int32_t hdr[6];
  std::string err;
  if (!rd(f, hdr, sizeof hdr) || (uint32_t)hdr[0] != kMagic || hdr[1] != 1) {
    err = "not an RXTB relline table: " + path;
  } else {
    t.n_a = hdr[2];
    t.n_mu0 = hdr[3];
    t.n_r = hdr[4];
    t.n_g = hdr[5];
    const size_t nn = t.node_count();
    bool ok = rvec(f, t.a, t.n_a) && rvec(f, t.mu0, t.n_mu0) && rvec(f, t.r, nn * t.n_r) &&
              rvec(f, t.gmin, nn * t.n_r) && rvec(f, t.gmax, nn * t.n_r) &&
              rvec(f, t.trff1, nn * t.n_r * t.n_g) && rvec(f, t.trff2, nn * t.n_r * t.n_g) &&
              rvec(f, t.cosne1, nn * t.n_r * t.n_g) && rvec(f, t.cosne2, nn * t.n_r * t.n_g);
    if (!ok) err = "truncated table file: " + path;
  }
  fclose(f);
  if (err.empty()) err = validate_reltable(t);
  return err;
}

std::string save_lptable(const LpTableHost &t, const std::string &path) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return "cannot open " + path;
  int32_t hdr[5] = {(int32_t)kMagic, 2, t.n_a, t.n_h, t.n_r};
  bool ok = wr(f, hdr, sizeof hdr) && wvec(f, t.a) && wvec(f, t.h) && wvec(f, t.rad) &&
            wvec(f, t.intens) && wvec(f, t.del) && wvec(f, t.del_inc);
  fclose(f);
  return ok ? "" : "write failed: " + path;
}

std::string load_lptable(LpTableHost &t, const std::string &path) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return "cannot open " + path;
  int32_t hdr[5];
  std::string err;
  if (!rd(f, hdr, sizeof hdr) || (uint32_t)hdr[0] != kMagic || hdr[1] != 2) {
    err = "not an RXTB lamp-post table: " + path;
  } else {
    t.n_a = hdr[2];
    t.n_h = hdr[3];
    t.n_r = hdr[4];
    const size_t n3 = (size_t)t.n_a * t.n_h * t.n_r;
    bool ok = rvec(f, t.a, t.n_a) && rvec(f, t.h, (size_t)t.n_a * t.n_h) &&
              rvec(f, t.rad, (size_t)t.n_a * t.n_r) && rvec(f, t.intens, n3) &&
              rvec(f, t.del, n3) && rvec(f, t.del_inc, n3);
    if (!ok) err = "truncated table file: " + path;
  }
  fclose(f);
  if (err.empty()) err = validate_lptable(t);
  return err;
}

}  // namespace rx
