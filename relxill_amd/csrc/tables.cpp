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
