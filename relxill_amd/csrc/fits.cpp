// fits.cpp -- loaders for the reference's FITS tables into the library's table
// layout (rx_tables.h):
//   relline table (rel_table_v0.5a.fits, src/reltable.c:121-304): binary-table
//     extensions "a" (column a) and "mu0" (column mu0) hold the axes; the node
//     of spin i and mu0 j is HDU number i * n_mu0 + j + 4 (1-based, :287-288),
//     with n_r rows of r, gmin, gmax and the n_g-vectors trff1, trff2, cosne1,
//     cosne2;
//   xillver tables (xillver-a-Ec5.fits, src/xilltable.c:169-277, :513-566) and
//     the returning-radiation table (table_returnRad_v20220301.fits,
//     src/Relreturn_Table.cpp:126-340) are read further below;
//   lamp-post table (rel_lp_table_v0.5b.fits, :313-448): extension "I_h", one
//     row per spin with columns a, hgrid[n_h], r[n_rad] and, for every height
//     k = 1..n_h, h<k>[n_rad] (intensity), del<k>, del_inc<k> (angles, stored
//     as their absolute values, :370-377).
// The dimensions are taken from the file (the reference compiles them in,
// src/common.h:69-81).  cfitsio is loaded at run time, so the library has no
// link-time dependency on it; without it these loaders return an error.
#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "rx_tables.h"

namespace rx {
namespace {

// The few cfitsio entry points used (fitsio.h: fftopn, ffclos, ffgerr, ffmahd,
// ffmnhd, ffgcno, ffgnrw, ffgtcl, ffgcv); the file handle stays opaque.
struct Fits {
  void *lib = nullptr;
  int (*open_table)(void **, const char *, int, int *) = nullptr;
  int (*close_file)(void *, int *) = nullptr;
  void (*get_errstatus)(int, char *) = nullptr;
  int (*movabs_hdu)(void *, int, int *, int *) = nullptr;
  int (*movnam_hdu)(void *, int, char *, int, int *) = nullptr;
  int (*get_colnum)(void *, int, char *, int *, int *) = nullptr;
  int (*get_num_rows)(void *, long *, int *) = nullptr;
  int (*get_coltype)(void *, int, int *, long *, long *, int *) = nullptr;
  int (*read_col)(void *, int, int, long long, long long, long long, void *, void *, int *, int *) = nullptr;
  std::string err;
};

constexpr int kReadOnly = 0, kBinaryTbl = 2, kCaseInsen = 0, kTFloat = 42, kTInt = 31, kTString = 16, kTDouble = 82;

// Candidate locations: RELXILL_CFITSIO_LIB, HEASoft's $HEADAS/lib (where an
// XSPEC installation keeps it), the loader's search path, $CONDA_PREFIX/lib.
Fits &fits_lib() {
  static Fits f;
  static std::once_flag once;
  std::call_once(once, [] {
    std::vector<std::string> cand;
    if (const char *p = getenv("RELXILL_CFITSIO_LIB")) cand.push_back(p);
    if (const char *p = getenv("HEADAS")) cand.push_back(std::string(p) + "/lib/libcfitsio.so");
    for (const char *n : {"libcfitsio.so", "libcfitsio.so.10", "libcfitsio.so.9", "libcfitsio.so.8"}) cand.push_back(n);
    if (const char *p = getenv("CONDA_PREFIX")) cand.push_back(std::string(p) + "/lib/libcfitsio.so");
    cand.push_back("/opt/conda/lib/libcfitsio.so");
    for (const std::string &c : cand) {
      f.lib = dlopen(c.c_str(), RTLD_NOW | RTLD_LOCAL);
      if (f.lib) break;
    }
    if (!f.lib) {
      f.err = "cfitsio not found (set RELXILL_CFITSIO_LIB to libcfitsio.so)";
      return;
    }
    auto sym = [&](const char *name) {
      void *s = dlsym(f.lib, name);
      if (!s && f.err.empty()) f.err = std::string("cfitsio lacks ") + name;
      return s;
    };
    f.open_table = (decltype(f.open_table))sym("fftopn");
    f.close_file = (decltype(f.close_file))sym("ffclos");
    f.get_errstatus = (decltype(f.get_errstatus))sym("ffgerr");
    f.movabs_hdu = (decltype(f.movabs_hdu))sym("ffmahd");
    f.movnam_hdu = (decltype(f.movnam_hdu))sym("ffmnhd");
    f.get_colnum = (decltype(f.get_colnum))sym("ffgcno");
    f.get_num_rows = (decltype(f.get_num_rows))sym("ffgnrw");
    f.get_coltype = (decltype(f.get_coltype))sym("ffgtcl");
    f.read_col = (decltype(f.read_col))sym("ffgcv");
  });
  return f;
}

// An open file; every call records the first failure in `err`.
struct FitsFile {
  Fits &f;
  void *fp = nullptr;
  int st = 0;
  std::string err;

  FitsFile(Fits &lib, const std::string &path) : f(lib) {
    if (f.open_table(&fp, path.c_str(), kReadOnly, &st)) fail("cannot open " + path);
  }
  ~FitsFile() {
    int s = 0;
    if (fp) f.close_file(fp, &s);
  }
  bool ok() const { return err.empty(); }
  void fail(const std::string &what) {
    if (!err.empty()) return;
    char msg[81] = {0};
    if (st) f.get_errstatus(st, msg);
    err = what + (st ? std::string(" (cfitsio: ") + msg + ")" : std::string());
  }
  void move_abs(int hdu) {
    int type = 0;
    if (ok() && f.movabs_hdu(fp, hdu, &type, &st)) fail("cannot move to HDU " + std::to_string(hdu));
  }
  void move_named(const char *name) {
    if (ok() && f.movnam_hdu(fp, kBinaryTbl, const_cast<char *>(name), 0, &st))
      fail(std::string("no extension ") + name);
  }
  int column(const std::string &name) {
    int c = 0;
    if (ok() && f.get_colnum(fp, kCaseInsen, const_cast<char *>(name.c_str()), &c, &st))
      fail("no column " + name);
    return c;
  }
  long rows() {
    long n = 0;
    if (ok() && f.get_num_rows(fp, &n, &st)) fail("cannot read the row count");
    return n;
  }
  long repeat(int col) {
    int type = 0;
    long rep = 0, width = 0;
    if (ok() && f.get_coltype(fp, col, &type, &rep, &width, &st)) fail("cannot read a column type");
    return rep;
  }
  // n float values of column col starting at (row, first element), 1-based
  void read(int col, long row, long first, long n, float *dst) {
    float nul = 0.0f;
    int anynul = 0;
    if (ok() && f.read_col(fp, kTFloat, col, row, first, n, &nul, dst, &anynul, &st)) fail("cannot read a column");
  }
  void read_dbl(int col, long row, long first, long n, double *dst) {
    double nul = 0.0;
    int anynul = 0;
    if (ok() && f.read_col(fp, kTDouble, col, row, first, n, &nul, dst, &anynul, &st)) fail("cannot read a column");
  }
  void read_int(int col, long row, long n, int *dst) {
    int nul = 0, anynul = 0;
    if (ok() && f.read_col(fp, kTInt, col, row, 1, n, &nul, dst, &anynul, &st)) fail("cannot read an integer column");
  }
  // one string cell (at most 71 characters), trailing blanks removed
  std::string read_str(int col, long row) {
    char buf[72] = {0};
    char *cell[1] = {buf};
    char nul[2] = " ";
    int anynul = 0;
    if (ok() && f.read_col(fp, kTString, col, row, 1, 1, nul, cell, &anynul, &st)) fail("cannot read a string column");
    std::string v(buf);
    while (!v.empty() && v.back() == ' ') v.pop_back();
    return v;
  }
};

// One axis column ("a" or "mu0") of a one-column extension.
std::vector<float> read_axis(FitsFile &ff, const char *ext, const char *col) {
  ff.move_named(ext);
  const int c = ff.column(col);
  const long n = ff.rows();
  std::vector<float> v(n > 0 ? n : 0);
  if (n > 0) ff.read(c, 1, 1, n, v.data());
  return v;
}

}  // namespace

std::string load_reltable_fits(RelTableHost &t, const std::string &path) {
  Fits &f = fits_lib();
  if (!f.err.empty()) return f.err;
  FitsFile ff(f, path);
  RelTableHost out;
  out.a = read_axis(ff, "a", "a");
  out.mu0 = read_axis(ff, "mu0", "mu0");
  out.n_a = (int)out.a.size();
  out.n_mu0 = (int)out.mu0.size();
  for (int ia = 0; ia < out.n_a && ff.ok(); ia++) {
    for (int im = 0; im < out.n_mu0 && ff.ok(); im++) {
      ff.move_abs(ia * out.n_mu0 + im + 4);
      const int cr = ff.column("r"), cgl = ff.column("gmin"), cgh = ff.column("gmax");
      const int ct1 = ff.column("trff1"), ct2 = ff.column("trff2");
      const int cc1 = ff.column("cosne1"), cc2 = ff.column("cosne2");
      const long nr = ff.rows(), ng = ff.repeat(ct1);
      if (!ff.ok()) break;
      if (ia == 0 && im == 0) {
        out.n_r = (int)nr;
        out.n_g = (int)ng;
        const size_t nn = out.node_count() * out.n_r;
        out.r.resize(nn);
        out.gmin.resize(nn);
        out.gmax.resize(nn);
        out.trff1.resize(nn * out.n_g);
        out.trff2.resize(nn * out.n_g);
        out.cosne1.resize(nn * out.n_g);
        out.cosne2.resize(nn * out.n_g);
      }
      if (nr != out.n_r || ng != out.n_g || ff.repeat(ct2) != ng || ff.repeat(cc1) != ng || ff.repeat(cc2) != ng) {
        ff.fail("node " + std::to_string(ia + 1) + "_" + std::to_string(im + 1) + " has other dimensions");
        break;
      }
      const size_t o = ((size_t)ia * out.n_mu0 + im) * out.n_r;
      ff.read(cr, 1, 1, nr, &out.r[o]);
      ff.read(cgl, 1, 1, nr, &out.gmin[o]);
      ff.read(cgh, 1, 1, nr, &out.gmax[o]);
      // a vector column is read row after row: row ir holds the n_g values of radius ir
      for (long ir = 0; ir < nr && ff.ok(); ir++) {
        const size_t q = (o + ir) * out.n_g;
        ff.read(ct1, ir + 1, 1, ng, &out.trff1[q]);
        ff.read(ct2, ir + 1, 1, ng, &out.trff2[q]);
        ff.read(cc1, ir + 1, 1, ng, &out.cosne1[q]);
        ff.read(cc2, ir + 1, 1, ng, &out.cosne2[q]);
      }
    }
  }
  if (!ff.ok()) return "relline table " + path + ": " + ff.err;
  std::string v = validate_reltable(out);
  if (!v.empty()) return "relline table " + path + ": " + v;
  t = std::move(out);
  return "";
}

std::string load_lptable_fits(LpTableHost &t, const std::string &path) {
  Fits &f = fits_lib();
  if (!f.err.empty()) return f.err;
  FitsFile ff(f, path);
  LpTableHost out;
  ff.move_named("I_h");
  const int ca = ff.column("a"), ch = ff.column("hgrid"), cr = ff.column("r");
  const long na = ff.rows();
  const long nh = ff.repeat(ch), nrad = ff.repeat(cr);
  if (ff.ok()) {
    out.n_a = (int)na;
    out.n_h = (int)nh;
    out.n_r = (int)nrad;
    out.a.resize(na);
    out.h.resize((size_t)na * nh);
    out.rad.resize((size_t)na * nrad);
    out.intens.resize((size_t)na * nh * nrad);
    out.del.resize((size_t)na * nh * nrad);
    out.del_inc.resize((size_t)na * nh * nrad);
    ff.read(ca, 1, 1, na, out.a.data());
  }
  std::vector<int> c_int(nh), c_del(nh), c_inc(nh);
  for (long k = 0; k < nh && ff.ok(); k++) {
    const std::string s = std::to_string(k + 1);
    c_int[k] = ff.column("h" + s);
    c_del[k] = ff.column("del" + s);
    c_inc[k] = ff.column("del_inc" + s);
    if (ff.ok() && (ff.repeat(c_int[k]) != nrad || ff.repeat(c_del[k]) != nrad || ff.repeat(c_inc[k]) != nrad))
      ff.fail("height column " + s + " has another length than r");
  }
  for (long ia = 0; ia < na && ff.ok(); ia++) {
    ff.read(ch, ia + 1, 1, nh, &out.h[(size_t)ia * nh]);
    ff.read(cr, ia + 1, 1, nrad, &out.rad[(size_t)ia * nrad]);
    for (long k = 0; k < nh && ff.ok(); k++) {
      const size_t o = ((size_t)ia * nh + k) * nrad;
      ff.read(c_int[k], ia + 1, 1, nrad, &out.intens[o]);
      ff.read(c_del[k], ia + 1, 1, nrad, &out.del[o]);
      ff.read(c_inc[k], ia + 1, 1, nrad, &out.del_inc[o]);
    }
  }
  if (!ff.ok()) return "lamp-post table " + path + ": " + ff.err;
  for (float &x : out.del) x = fabsf(x);
  for (float &x : out.del_inc) x = fabsf(x);
  std::string v = validate_lptable(out);
  if (!v.empty()) return "lamp-post table " + path + ": " + v;
  t = std::move(out);
  return "";
}

std::string load_rradtable_fits(RradTableHost &t, const std::string &path) {
  Fits &f = fits_lib();
  if (!f.err.empty()) return f.err;
  FitsFile ff(f, path);
  RradTableHost out;
  // SPIN: the spins (get_reltable_axis_double, src/Relreturn_Table.cpp:126-150)
  ff.move_named("SPIN");
  const int ca = ff.column("a");
  const long ns = ff.rows();
  if (ff.ok()) {
    out.n_spin = (int)ns;
    out.spin.resize(ns);
    ff.read_dbl(ca, 1, 1, ns, out.spin.data());
  }
  // FRAC<k>: one extension per spin (:262-300, extension names :331)
  for (long k = 0; k < ns && ff.ok(); k++) {
    char ext[16];
    snprintf(ext, sizeof ext, "FRAC%02ld", k + 1);
    ff.move_named(ext);
    const int c_rlo = ff.column("rlo"), c_rhi = ff.column("rhi"), c_tf = ff.column("tf_r");
    const int c_gmin = ff.column("gmin"), c_gmax = ff.column("gmax"), c_fg = ff.column("frac_g");
    const int c_fret = ff.column("f_ret");
    const long nr = ff.rows();
    if (!ff.ok()) break;
    if (k == 0) {
      out.n_rad = (int)nr;
      out.n_g = nr > 0 ? (int)(ff.repeat(c_fg) / nr) : 0;
      const size_t nr2 = (size_t)nr * nr;
      out.rlo.resize(ns * nr);
      out.rhi.resize(ns * nr);
      out.f_ret.resize(ns * nr);
      out.tf_r.resize(ns * nr2);
      out.gmin.resize(ns * nr2);
      out.gmax.resize(ns * nr2);
      out.frac_g.resize(ns * nr2 * out.n_g);
    }
    if (nr != out.n_rad || ff.repeat(c_tf) != nr || ff.repeat(c_fg) != nr * out.n_g || out.n_g < 1) {
      ff.fail(std::string(ext) + " has other dimensions");
      break;
    }
    const size_t o1 = (size_t)k * nr, o2 = (size_t)k * nr * nr;
    ff.read_dbl(c_rlo, 1, 1, nr, &out.rlo[o1]);
    ff.read_dbl(c_rhi, 1, 1, nr, &out.rhi[o1]);
    ff.read_dbl(c_fret, 1, 1, nr, &out.f_ret[o1]);
    for (long i = 0; i < nr && ff.ok(); i++) {  // row i: incident zone i
      ff.read_dbl(c_tf, i + 1, 1, nr, &out.tf_r[o2 + (size_t)i * nr]);
      ff.read_dbl(c_gmin, i + 1, 1, nr, &out.gmin[o2 + (size_t)i * nr]);
      ff.read_dbl(c_gmax, i + 1, 1, nr, &out.gmax[o2 + (size_t)i * nr]);
      ff.read_dbl(c_fg, i + 1, 1, nr * out.n_g, &out.frac_g[(o2 + (size_t)i * nr) * out.n_g]);
    }
  }
  if (!ff.ok()) return "returning-radiation table " + path + ": " + ff.err;
  std::string v = validate_rradtable(out);
  if (!v.empty()) return "returning-radiation table " + path + ": " + v;
  t = std::move(out);
  return "";
}

// Parameter names of xillver tables and their ids (src/common.h:141-161,
// src/xilltable.c:24-27; kTe shares the id of Ecut, A_CO that of A_Fe).
static int xill_param_id(const std::string &name) {
  static const struct {
    const char *name;
    int id;
  } kNames[] = {{"Gamma", XP_GAM}, {"A_Fe", XP_AFE}, {"logXi", XP_LXI}, {"Ecut", XP_ECT}, {"kTe", XP_ECT},
                {"Dens", XP_DNS}, {"kTbb", XP_KTB}, {"A_CO", XP_AFE}, {"Frac", XP_FRA}, {"Incl", XP_INC}};
  for (const auto &n : kNames)
    if (name == n.name) return n.id;
  return -1;
}

std::string load_xilltable_fits(XillTableHost &t, const std::string &path) {
  Fits &f = fits_lib();
  if (!f.err.empty()) return f.err;
  FitsFile ff(f, path);
  XillTableHost out;
  // PARAMETERS: one row per parameter; the column numbers the reference reads
  // (src/xilltable.c:182-183, :205-206, :219-220)
  ff.move_named("PARAMETERS");
  const long n_par = ff.rows();
  if (ff.ok() && n_par != 5 && n_par != 6)  // (new_xillTable, src/xilltable.c:78-81)
    return "xillver table " + path + ": tables have five or six parameters";
  out.n_par = (int)n_par;
  if (ff.ok()) ff.read_int(9, 1, n_par, out.n_val);
  for (int k = 0; k < out.n_par && ff.ok(); k++) {
    const std::string name = ff.read_str(1, k + 1);
    out.pid[k] = xill_param_id(name);
    if (ff.ok() && out.pid[k] < 0) ff.fail("unknown parameter " + name);
    if (ff.ok() && (out.n_val[k] < 2 || out.n_val[k] > ff.repeat(10))) ff.fail("parameter " + name + " has a bad NUMBVALS");
    if (!ff.ok()) break;
    out.val[k].resize(out.n_val[k]);
    ff.read(10, k + 1, 1, out.n_val[k], out.val[k].data());
  }
  // ENERGIES: ENERG_LO, ENERG_HI (src/xilltable.c:240-276)
  ff.move_named("ENERGIES");
  const long n_e = ff.rows();
  if (ff.ok()) {
    out.n_ener = (int)n_e;
    out.elo.resize(n_e);
    out.ehi.resize(n_e);
    ff.read(1, 1, 1, n_e, out.elo.data());
    ff.read(2, 1, 1, n_e, out.ehi.data());
  }
  // SPECTRA: row r + 1 holds spectrum r in column 2 (src/xilltable.c:532-545)
  ff.move_named("SPECTRA");
  if (ff.ok()) {
    const size_t rows = out.n_rows();
    if ((size_t)ff.rows() < rows || ff.repeat(2) < n_e) {
      ff.fail("SPECTRA has fewer rows or bins than the parameter grid needs");
    } else {
      out.spec.resize(rows * (size_t)n_e);
      for (size_t r = 0; r < rows && ff.ok(); r++) ff.read(2, (long)r + 1, 1, n_e, &out.spec[r * n_e]);
    }
  }
  if (!ff.ok()) return "xillver table " + path + ": " + ff.err;
  std::string v = validate_xilltable(out);
  if (!v.empty()) return "xillver table " + path + ": " + v;
  t = std::move(out);
  return "";
}

}  // namespace rx
