// rx_tables.h -- host-side relativistic tables: layout, synthetic generator,
// on-disk format.
//
// Layout mirrors the reference's in-memory tables (src/reltable.h:24-69) with
// run-time dimensions and flat float32 arrays:
//   relline table  a[n_a], mu0[n_mu0], node-major r/gmin/gmax[n_a][n_mu0][n_r],
//                  trff1/trff2/cosne1/cosne2[n_a][n_mu0][n_r][n_g]
//   lamp-post      a[n_a], h[n_a][n_h], rad[n_a][n_r],
//                  intens/del/del_inc[n_a][n_h][n_r]
// The real FITS tables (rel_table_v0.5a.fits, rel_lp_table_v0.5b.fits) are not
// available here; the synthetic generator builds smooth, seeded fields with the
// grid invariants the algorithm relies on (ascending a/mu0, descending radii
// ending at the ISCO, 1 <= r <= 1000, gmin < gmax, trff > 0, 0 < cosne <= 1,
// per-spin ascending h grids, ascending LP radii ending at 1000).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace rx {

struct RelTableHost {
  int n_a = 0, n_mu0 = 0, n_r = 0, n_g = 0;
  std::vector<float> a, mu0, r, gmin, gmax, trff1, trff2, cosne1, cosne2;
  size_t node_count() const { return (size_t)n_a * n_mu0; }
};

struct LpTableHost {
  int n_a = 0, n_h = 0, n_r = 0;
  std::vector<float> a, h, rad, intens, del, del_inc;
};

double kerr_rms(double a);
double kerr_rplus(double a);

RelTableHost make_synthetic_reltable(uint64_t seed, int n_a, int n_mu0, int n_r, int n_g);
LpTableHost make_synthetic_lptable(uint64_t seed, int n_a, int n_h, int n_r);

// Own binary format ("RXTB"): returns an empty string on success, else the error.
std::string save_reltable(const RelTableHost &t, const std::string &path);
std::string load_reltable(RelTableHost &t, const std::string &path);
std::string save_lptable(const LpTableHost &t, const std::string &path);
std::string load_lptable(LpTableHost &t, const std::string &path);

// The reference's FITS tables (rel_table_v0.5a.fits / rel_lp_table_v0.5b.fits
// layout, dimensions from the file); cfitsio is loaded at run time (fits.cpp).
std::string load_reltable_fits(RelTableHost &t, const std::string &path);
std::string load_lptable_fits(LpTableHost &t, const std::string &path);

// Checks the grid invariants; returns an empty string when they hold.
std::string validate_reltable(const RelTableHost &t);
std::string validate_lptable(const LpTableHost &t);

// xillver reflection table (the xillTable of src/xilltable.h): n_par = 5 or 6
// parameters, the last one the inclination; parameter k has the id pid[k]
// (PARAM_* of src/common.h:141-149) and n_val[k] ascending values in val[k];
// the energy grid has n_ener bins [elo, ehi]; spec holds row
// r = ((((i0 n1 + i1) n2 + i2) n3 + i3) n4 + i4) (five parameters; one more
// index, outermost, for six) at spec[r * n_ener] (row number r + 1 of the
// SPECTRA extension, src/xilltable.c:326-337), as stored in the file (not yet
// divided by 10^logXi).  A sixth parameter is the first one: it is
// interpolated outside the other five (interp_6d_tab_incl, :997-1045).
enum XillParamId : int {
  XP_GAM = 0, XP_AFE = 1, XP_LXI = 2, XP_ECT = 3, XP_DNS = 4, XP_KTB = 5, XP_FRA = 6, XP_INC = 7
};
constexpr int XILL_NPAR = 6;  // parameters of a table at most

struct XillTableHost {
  int n_par = 0;
  int pid[XILL_NPAR] = {};
  int n_val[XILL_NPAR] = {};
  std::vector<float> val[XILL_NPAR];
  int n_ener = 0;
  std::vector<float> elo, ehi;
  std::vector<float> spec;
  size_t n_rows() const {
    size_t n = 1;
    for (int k = 0; k < n_par; k++) n *= (size_t)n_val[k];
    return n;
  }
  int n_incl() const { return n_par ? n_val[n_par - 1] : 0; }
  int dim_of(int id) const {
    for (int k = 0; k < n_par; k++)
      if (pid[k] == id) return k;
    return -1;
  }
};

// Returning-radiation table (returnTable / tabulatedReturnFractions,
// src/Relreturn_Table.h; table_returnRad_v20220301.fits): per spin (ascending)
// n_rad radial zones [rlo, rhi] (ascending, the first starting at the spin's
// ISCO) and, for every pair (incident zone ro, emitting zone re), the fraction
// tf_r[ro][re] of the radiation emitted at re that returns to ro, the g range
// [gmin, gmax][ro][re] and the distribution frac_g[ro][re][n_g] of the energy
// shift over n_g bins of that range.  Flat, spin-major arrays.
struct RradTableHost {
  int n_spin = 0, n_rad = 0, n_g = 0;
  std::vector<double> spin;
  std::vector<double> rlo, rhi;                  // [spin][rad]
  std::vector<double> tf_r, gmin, gmax;          // [spin][ro][re]
  std::vector<double> frac_g;                    // [spin][ro][re][g]
  std::vector<double> f_ret;                     // [spin][rad]
};
RradTableHost make_synthetic_rradtable(uint64_t seed, int n_spin, int n_rad, int n_g);
std::string validate_rradtable(const RradTableHost &t);
// layout of src/Relreturn_Table.cpp:126-300: extension SPIN (column a), then
// FRAC01, FRAC02, ... (one per spin) with rlo, rhi, tf_r, gmin, gmax, frac_g, f_ret
std::string load_rradtable_fits(RradTableHost &t, const std::string &path);

// Seeded synthetic table with the parameters Gamma, A_Fe, logXi, Ecut (or
// kTe: the same parameter id), Incl (in that order) and n_ener log bins over
// 0.07-1000 keV; the fourth axis is log-spaced over [e4_lo, e4_hi] (Ecut
// 5-1000 keV; kTe 1-400 keV in the nthcomp tables).
XillTableHost make_synthetic_xilltable(uint64_t seed, int n_gam, int n_afe, int n_lxi, int n_ect,
                                       int n_incl, int n_ener, double e4_lo = 5.0, double e4_hi = 1000.0);
// Seeded synthetic table of the blackbody models with the parameters kTbb
// (0.5-10 keV), A_Fe, logXi (1-4.7), Dens (15-19), Incl, as xillverNS-2.fits
// names them, and n_ener log bins over 0.07-1000 keV.
XillTableHost make_synthetic_xilltable_ns(uint64_t seed, int n_ktbb, int n_afe, int n_lxi, int n_dens, int n_incl,
                                          int n_ener);
// Seeded synthetic six-parameter table: Gamma, A_Fe, logXi, Ecut (kTe), Dens
// (15-19), Incl.
XillTableHost make_synthetic_xilltable_dens(uint64_t seed, int n_dens, int n_gam, int n_afe, int n_lxi, int n_ect,
                                            int n_incl, int n_ener, double e4_lo = 5.0, double e4_hi = 1000.0);
std::string validate_xilltable(const XillTableHost &t);
// The spectra divided by 10^logXi of their row and, for a table with a
// density axis, by 10^(logN - 15) (renorm_xill_spec, src/xilltable.c:478-511,
// applied when a spectrum is loaded, :513-566).
std::vector<float> xilltable_renormed(const XillTableHost &t);
// xillver-a-Ec5.fits layout: extensions PARAMETERS (NAME in column 1,
// NUMBVALS in column 9, VALUE in column 10), ENERGIES (ENERG_LO, ENERG_HI),
// SPECTRA (spectrum in column 2), src/xilltable.c:169-277, :513-566.
std::string load_xilltable_fits(XillTableHost &t, const std::string &path);

}  // namespace rx
