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

}  // namespace rx
