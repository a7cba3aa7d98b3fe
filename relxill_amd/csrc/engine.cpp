// engine.cpp -- host runtime: parameters, tables on the device, batch planning
// and kernel launches.
#include "engine.h"
#include "rx_math.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace rx {

namespace {
constexpr double kPi = 3.14159265358979323846;
thread_local int t_status = 0;
thread_local std::string t_error;
thread_local const char *t_error_static = nullptr;  // set instead of t_error where nothing may allocate
}  // namespace

void set_last(int status, const std::string &msg) {
  t_status = status;
  t_error = msg;
  t_error_static = nullptr;
}
void set_last_static(int status, const char *msg) {
  t_status = status;
  t_error_static = msg ? msg : "";
}
int last_status() { return t_status; }
const char *last_error() { return t_error_static ? t_error_static : t_error.c_str(); }

Global &G() {
  static Global g;
  return g;
}

DevBuf::~DevBuf() {
  if (p) (void)hipFree(p);
}

int DevBuf::ensure(size_t n) {
  if (n == 0) n = 8;
  if (p && bytes >= n) return 0;
  // a buffer that has to grow grows by half again at least: hipFree waits for
  // the whole device, so a reused buffer (a server slot's) should rarely grow
  if (p) n = std::max(n, bytes + bytes / 2);
  n = (n + 255) & ~(size_t)255;
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) {
    p = nullptr;
    return (int)e;
  }
  bytes = n;
  return 0;
}

bool phys_norm_env() {
  // do_not_normalize_relline, src/relutility.c:386-396
  const char *env = getenv("RELLINE_PHYSICAL_NORM");
  return env && (int)strtod(env, nullptr) == 1;
}

int n_model_params(int model) {
  switch (model) {
    case MOD_RELLINE: return 10;
    case MOD_RELCONV: return 8;
    case MOD_RELLINELP: return 10;
    case MOD_RELCONVLP: return 9;
    case MOD_RELXILL_CONV: return 8;
    case MOD_RELXILL: return 13;
    case MOD_RELXILLLP: return 14;
    case MOD_RELXILLCP: return 14;
    case MOD_RELXILLLPCP: return 17;
    case MOD_RELXILLNS: return 13;
    case MOD_XILLVER: return 7;
    case MOD_XILLVERCP: return 8;
    case MOD_XILLVERNS: return 7;
    default: return -1;
  }
}

// get_rel_params (src/ModelDefinition.cpp:278-334) with check_parameter_bounds (:168-273)
int parse_request(int model, const double *par, bool phys_norm, ReqParams &p, std::string &err, int prim) {
  memset(&p, 0, sizeof(p));
  p.model = model;
  p.prim = prim;
  p.nth_off = -1;
  p.nth_sol = -1;
  p.pgrid_off = -1;
  p.prof_off = -1;
  double incl_deg = 0, htop = 0;
  int return_rad = 0;
  switch (model) {
    case MOD_RELLINE:  // lineE Index1 Index2 Rbr a Incl Rin Rout z limb
      p.emis_type = EMIS_BKN;
      p.lineE = par[0]; p.emis1 = par[1]; p.emis2 = par[2]; p.rbr = par[3]; p.a = par[4];
      incl_deg = par[5]; p.rin = par[6]; p.rout = par[7]; p.z = par[8];
      p.limb = (int)lround(par[9]);
      break;
    case MOD_RELXILL:  // Index1 Index2 Rbr a Incl Rin Rout z gamma logxi Afe Ecut refl_frac
      p.emis_type = EMIS_BKN;
      p.emis1 = par[0]; p.emis2 = par[1]; p.rbr = par[2]; p.a = par[3]; incl_deg = par[4];
      p.rin = par[5]; p.rout = par[6]; p.z = par[7]; p.gamma = par[8];
      // get_xill_params (src/ModelDefinition.cpp:339-390); the cut-off power
      // law of the broken power-law model keeps Ecut as given
      p.xgam = par[8];
      p.xect = par[11];
      p.refl_frac = par[12];
      break;
    case MOD_RELXILLLP:  // h beta a Incl Rin Rout z gamma logxi Afe Ecut refl_frac
                         // switch_returnrad switch_reflfrac_boost
      p.emis_type = EMIS_LP;
      p.height = par[0]; p.beta = par[1]; p.a = par[2]; incl_deg = par[3]; p.rin = par[4];
      p.rout = par[5]; p.z = par[6]; p.gamma = par[7];
      return_rad = (int)lround(par[12]);
      p.xgam = par[7];
      p.xect = par[10];
      p.refl_frac = par[11];
      p.xboost = (int)lround(par[13]);
      break;
    case MOD_XILLVER:  // gamma Afe Ecut logxi z Incl refl_frac (xillverCp and xillverNS are
                       // remapped to this order): no relativistic parameters and no
                       // checks beyond the table's clamps (LocalModel::xillver_model,
                       // src/LocalModel.cpp:104-128, reads only get_xill_params)
      p.xgam = par[0];
      p.xect = par[2];
      p.xect_src = par[2];
      p.z = par[4];
      p.incl = par[5] * kPi / 180;
      p.refl_frac = par[6];
      p.norm = 0.5 * cos(par[5] * kPi / 180);  // norm_factor_semi_infinite_slab, src/Xillspec.cpp:528-530
      p.n_zones = 1;
      p.g_so = 1.0;
      return RX_OK;
    case MOD_RELCONV:  // Index1 Index2 Rbr a Incl Rin Rout limb
    case MOD_RELXILL_CONV:
      p.emis_type = EMIS_BKN;
      p.emis1 = par[0]; p.emis2 = par[1]; p.rbr = par[2]; p.a = par[3]; incl_deg = par[4];
      p.rin = par[5]; p.rout = par[6]; p.limb = (int)lround(par[7]);
      break;
    case MOD_RELLINELP:  // lineE h a Incl Rin Rout z limb gamma switch_returnrad
      p.emis_type = EMIS_LP;
      p.lineE = par[0]; p.height = par[1]; p.a = par[2]; incl_deg = par[3]; p.rin = par[4];
      p.rout = par[5]; p.z = par[6]; p.limb = (int)lround(par[7]); p.gamma = par[8];
      return_rad = (int)lround(par[9]);
      break;
    case MOD_RELCONVLP:  // h beta a Incl Rin Rout limb gamma switch_returnrad
      p.emis_type = EMIS_LP;
      p.height = par[0]; p.beta = par[1]; p.a = par[2]; incl_deg = par[3]; p.rin = par[4];
      p.rout = par[5]; p.limb = (int)lround(par[6]); p.gamma = par[7];
      return_rad = (int)lround(par[8]);
      break;
    default:
      err = "unknown model id";
      return RX_ERR_ARG;
  }
  p.incl = incl_deg * kPi / 180;
  // do_renorm_model (src/relutility.c:603-623)
  if (model < 0) {
    p.do_renorm = (p.emis_type == EMIS_LP || phys_norm) ? 0 : 1;
  } else {
    p.do_renorm = phys_norm ? 0 : 1;
  }
  // check_parameter_bounds
  int status = RX_OK;
  if (p.rin < 0) p.rin = -1.0 * p.rin * kerr_rms(p.a);
  if (p.rout < 0) p.rout = -1.0 * p.rout * kerr_rms(p.a);
  if (p.rbr < 0) p.rbr = -1.0 * p.rbr * kerr_rms(p.a);
  if (p.rout <= p.rin) status = RX_ERR_PARAM;
  const double rms = kerr_rms(p.a);
  p.rms = rms;  // the device uses this value, not its own pow/sqrt rounding
  if (p.rin < rms) p.rin = rms;
  if (p.a > 0.9982) { err = "spin a > 0.9982"; return RX_ERR_PARAM; }
  if (p.a < -1) { err = "spin a < -1"; return RX_ERR_PARAM; }
  if (p.incl < 3 * kPi / 180 || p.incl > 87 * kPi / 180) {
    err = "inclination outside 3-87 deg";
    return RX_ERR_PARAM;
  }
  if (p.rout <= p.rin) { err = "Rout <= Rin"; return RX_ERR_PARAM; }
  if (p.rout > 1000.0) p.rout = 1000.0;
  if (p.emis_type == EMIS_BKN) {
    if (p.rbr < p.rin) p.rbr = p.rin;
    if (p.rbr > p.rout) p.rbr = p.rout;
  }
  if (p.emis_type == EMIS_LP) {
    if (p.beta < 0) p.beta = 0.0;
    if (p.beta > 0.99) p.beta = 0.99;
    if (p.height < 0) p.height = -1.0 * p.height * kerr_rplus(p.a);
    if (htop < 0) htop = -1.0 * htop * kerr_rplus(p.a);
    const double r_event = kerr_rplus(p.a);
    if ((1.1 * r_event - p.height) > 1e-4) p.height = r_event * 1.1;
  }
  if (status != RX_OK) {
    err = "Rin >= Rout";
    return status;
  }
  if (return_rad != 0) {
    // add_returnrad_emis (src/Rellp.cpp:483-507): 1 adds the returning
    // radiation, -1 and 2 replace the direct emissivity by it; relxilllp
    // corrects it with its zones' xillver factors above spin 0
    // (SPIN_MIN_RRAD_CALC_CORRFAC, src/Relxill.h:32, src/Relxill.cpp:337-341)
    if (!(return_rad == 1 || return_rad == -1 || return_rad == 2)) {
      err = "switch_returnrad must be 0, 1, -1 or 2";
      return RX_ERR_PARAM;
    }
  }
  p.return_rad = return_rad;
  p.rr_corr = (model == MOD_RELXILLLP && return_rad != 0 && p.a > 0.0) ? 1 : 0;
  p.norm = 1.0;
  if (model < 0 && p.emis_type == EMIS_BKN) {
    // norm_factor_semi_infinite_slab, src/Xillspec.cpp:528-530
    p.norm = 0.5 * cos((p.incl * 180.0 / kPi) * kPi / 180);
  }
  p.n_zones = 1;
  p.g_so = 1.0;
  if (model == MOD_RELXILL || model == MOD_RELXILLLP) {
    // zones (get_num_zones, src/relutility.c:506-544: N_ZONES for the lamp
    // post, one otherwise) and their edges
    p.n_zones = (model == MOD_RELXILLLP) ? LP_ZONES : 1;
    rzone_grid(p.rin, p.rout, p.n_zones, p.height, p.zgrid);
    if (p.emis_type == EMIS_LP) {
      // energy_shift_source_obs (src/Relphysics.cpp:239-254, :134-136, :217-230)
      const double g_inf_0 = sqrt(1.0 - (2 * p.height / (p.height * p.height + p.a * p.a)));
      p.g_so = g_inf_0;
      if (!(p.beta < 1e-4)) p.g_so = g_inf_0 * (sqrt(1.0 - p.beta * p.beta) / (1.0 + p.beta * cos(kPi - p.incl)));
    }
    // the lamp post gives Ecut in the observer frame; the xillver tables and
    // the normalisation use the source's (src/Relxill.cpp:197-200).  kTe of
    // the nthcomp models is already the source's.
    p.xect_src = (p.emis_type == EMIS_LP && prim == PRIM_CUTOFFPL) ? p.xect / p.g_so : p.xect;
  }
  return RX_OK;
}

void rzone_grid(double rmin, double rmax, int nzones, double h, double *rgrid) {
  // RadialGrid::calculate_radial_grid (src/IonGradient.cpp:457-489): log
  // spacing below the source height, 1/r spacing above
  if (nzones == 1) {
    rgrid[0] = rmin;
    rgrid[1] = rmax;
    return;
  }
  double r_transition = rmin;
  int indr = 0;
  if (h > rmin) {
    r_transition = h;
    for (int ii = 0; ii < nzones + 1; ii++) {  // get_log_grid, src/relutility.c:399-405
      rgrid[ii] = exp(1.0 * ii / (nzones + 1 - 1) * (log(rmax) - log(rmin)) + log(rmin));
    }
    indr = bsearch_asc(rgrid, nzones + 1, r_transition);
    r_transition = rgrid[indr];
  }
  if (indr < nzones) {
    const double rlo = r_transition, rhi = rmax;
    for (int ii = indr; ii < nzones + 1; ii++) {
      rgrid[ii] = 1.0 * (ii - indr) / (nzones - indr) * (1.0 / rhi - 1.0 / rlo) + 1.0 / rlo;
      rgrid[ii] = fabs(1.0 / rgrid[ii]);
    }
  }
}

// binary_search_float (src/relutility.c:135-151)
static int bsearch_float(const float *arr, int n, float val) {
  if (n <= 1) return -1;
  int klo = 0, khi = n - 1;
  while (khi - klo > 1) {
    const int k = (khi + klo) / 2;
    if (arr[k] > val) khi = k;
    else klo = k;
  }
  return klo;
}

void xill_corner(const XillTableHost &t, double gam, double afe, double lxi, double ect, double dens, double ktbb,
                 double incl_deg, int *ind, double *fac, bool clamp_incl, int *inc_ind, double *inc_fac, int *out_ind,
                 double *out_fac) {
  // the input value of every table parameter as float (get_xilltab_paramvals,
  // src/xilltable.c:278-295); kTbb is 0 but for the blackbody model, Frac 0
  float v[8] = {(float)gam, (float)afe, (float)lxi, (float)ect, (float)dens, (float)ktbb, 0.0f, (float)incl_deg};
  int k_ind[XILL_NPAR];
  for (int k = 0; k < t.n_par; k++) {
    int i = bsearch_float(t.val[k].data(), t.n_val[k], v[t.pid[k]]);
    if (i < 0) i = 0;
    else if (i > t.n_val[k] - 2) i = t.n_val[k] - 2;
    k_ind[k] = i;
  }
  // factors in float after clamping the value into the table (relxill
  // models leave the inclination alone: it is not interpolated)
  double k_fac[XILL_NPAR];
  for (int k = 0; k < t.n_par; k++) {
    float &x = v[t.pid[k]];
    const std::vector<float> &vals = t.val[k];
    if (t.pid[k] != XP_INC || clamp_incl) {
      if (x < vals[0]) x = vals[0];
      else if (x > vals[t.n_val[k] - 1]) x = vals[t.n_val[k] - 1];
    }
    const float num = x - vals[k_ind[k]];
    const float den = vals[k_ind[k] + 1] - vals[k_ind[k]];
    k_fac[k] = (float)(num / den);
  }
  // ensure_ecut_within_boundarys (src/xilltable.c:1054-1067) reads the axis
  // numbered by the id of Ecut (3): Ecut's own axis in tables that list it
  // fourth, as xillver-a-Ec5.fits does
  const int k_ect = t.dim_of(XP_ECT);
  if (k_ect >= 0 && t.pid[k_ect] < t.n_par) {
    const std::vector<float> &vals = t.val[t.pid[k_ect]];
    if (ect <= vals[0]) k_fac[k_ect] = 0.0;
    if (ect >= vals[vals.size() - 1]) k_fac[k_ect] = 1.0;
  }
  // the four inner axes (those after the sixth parameter of a six-parameter
  // table, which is interpolated outside them: interp_6d_tab_incl,
  // src/xilltable.c:997-1045), then that outer one and the inclination
  const int off = t.n_par - 5;
  for (int k = 0; k < 4; k++) {
    ind[k] = k_ind[k + off];
    fac[k] = k_fac[k + off];
  }
  if (out_ind) *out_ind = off ? k_ind[0] : 0;
  if (out_fac) *out_fac = off ? k_fac[0] : 0.0;
  if (inc_ind) *inc_ind = k_ind[t.n_par - 1];
  if (inc_fac) *inc_fac = k_fac[t.n_par - 1];
}

// ---------------------------------------------------------------------------
// tables and grids
// ---------------------------------------------------------------------------

// Tables named by RELXILL_TABLE_PATH.  The reference reads its tables only
// from that directory (get_relxill_table_path, src/relutility.c:320-328) and
// stops with an error when it cannot open one (src/reltable.c:255-268), so a
// set path never falls back to synthetic tables here: the first of the file
// names that exists is loaded, and a missing or unreadable table leaves its
// error in *err (returned by every request that needs the table).  Synthetic
// tables are only built when the variable is unset.
static bool file_exists(const std::string &path) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return false;
  fclose(f);
  return true;
}

template <class T, class LoadFits, class LoadOwn>
static bool load_from_path(T &dst, const char *fits_name, const char *own_name, LoadFits load_fits,
                           LoadOwn load_own, std::string &err) {
  const std::string dir = getenv("RELXILL_TABLE_PATH");
  const std::string fits = dir + "/" + fits_name;
  const std::string own = own_name ? dir + "/" + own_name : std::string();
  T t;
  std::string e;
  if (file_exists(fits)) {
    e = load_fits(t, fits);
    if (!e.empty()) e = fits + ": " + e;
  } else if (own_name && file_exists(own)) {
    e = load_own(t, own);
    if (!e.empty()) e = own + ": " + e;
  } else {
    e = std::string("RELXILL_TABLE_PATH=") + dir + " holds no " + fits_name +
        (own_name ? std::string(" (nor ") + own_name + ")" : std::string());
  }
  if (!e.empty()) {
    err = e;
    return false;
  }
  dst = std::move(t);
  err.clear();
  return true;
}

void ensure_tables_host() {
  Global &g = G();
  const bool from_path = getenv("RELXILL_TABLE_PATH") != nullptr;
  if (!g.rel_loaded) {
    // the reference's table file and name first (src/common.h:73), then our own format
    if (!from_path) {
      g.rel = make_synthetic_reltable(1, 25, 22, 100, 20);
      g.rel_err.clear();
      g.rel_loaded = true;
    } else {
      g.rel_loaded = load_from_path(g.rel, "rel_table_v0.5a.fits", "rel_table.rxtb", load_reltable_fits,
                                    load_reltable, g.rel_err);
    }
    if (g.rel_loaded) g.rel_dirty = true;
  }
  if (g.rel_loaded && g.rel_dirty) {
    g.rel_gmin_all = *std::min_element(g.rel.gmin.begin(), g.rel.gmin.end());
    g.rel_gmax_all = *std::max_element(g.rel.gmax.begin(), g.rel.gmax.end());
  }
  if (!g.lp_loaded) {
    if (!from_path) {
      g.lp = make_synthetic_lptable(3, 20, 100, 100);
      g.lp_err.clear();
      g.lp_loaded = true;
    } else {
      g.lp_loaded = load_from_path(g.lp, "rel_lp_table_v0.5b.fits", "rel_lp_table.rxtb", load_lptable_fits,
                                   load_lptable, g.lp_err);
    }
    if (g.lp_loaded) g.lp_dirty = true;
  }
  if (g.conv_ener.empty()) g.conv_dirty = true;
  if (g.coarse_ener.empty()) {
    // get_coarse_xillver_energrid (src/Xillspec.cpp:164-177) via get_log_grid
    const int n = N_ENER_COARSE;
    g.coarse_ener.resize(n + 1);
    for (int ii = 0; ii < n + 1; ii++) {
      double e = 1.0 * ii / (n + 1 - 1) * (log(1000.0) - log(0.1)) + log(0.1);
      g.coarse_ener[ii] = exp(e);
    }
    g.coarse_ln.resize(n);
    for (int ii = 0; ii < n; ii++) g.coarse_ln[ii] = log(0.5 * (g.coarse_ener[ii] + g.coarse_ener[ii + 1]));
  }
  if (g.conv_dirty) {
    const int n = g.conv_n;
    g.conv_ener.resize(n + 1);
    // get_log_grid, src/relutility.c:399-405
    for (int ii = 0; ii < n + 1; ii++) {
      double e = 1.0 * ii / (n + 1 - 1) * (log(g.conv_emax) - log(g.conv_emin)) + log(g.conv_emin);
      g.conv_ener[ii] = exp(e);
    }
    g.conv_cfac.resize(n);
    g.conv_cfinv.resize(n);
    g.conv_inv_emid.resize(n);
    g.conv_ln_emid.resize(n);
    for (int ii = 0; ii < n; ii++) {
      // calculate_energyflux_conversion, src/Relbase.cpp:93-103
      const double emid = 0.5 * (g.conv_ener[ii] + g.conv_ener[ii + 1]);
      const double de = g.conv_ener[ii + 1] - g.conv_ener[ii];
      g.conv_cfac[ii] = emid / de;
      g.conv_cfinv[ii] = de / emid;
      g.conv_inv_emid[ii] = 1.0 / emid;
      g.conv_ln_emid[ii] = log(emid);
    }
    g.conv_i1kev = bsearch_asc(g.conv_ener.data(), n + 1, 1.0);
    g.conv_band_lo = n;
    g.conv_band_hi = 0;
    for (int ii = 0; ii < n; ii++) {
      if (g.conv_ener[ii] >= 0.01 && g.conv_ener[ii + 1] < 1000.0) {
        g.conv_band_lo = std::min(g.conv_band_lo, ii);
        g.conv_band_hi = std::max(g.conv_band_hi, ii + 1);
      }
    }
    // per-pass twiddles of the device FFTs (forward n points, inverse n/2):
    // pass (Ns, R) stores exp(-2 pi i r k / (Ns R)) at [r][k], k < Ns
    for (int which = 0; which < 2; which++) {
      std::vector<double> &t = which == 0 ? g.conv_twp_fwd : g.conv_twp_inv;
      const int N = which == 0 ? n : n / 2;
      t.clear();
      for (int Ns = 1; Ns < N;) {
        const int left = N / Ns;
        const int R = left % 8 == 0 ? 8 : left % 4 == 0 ? 4 : 2;  // the device plan (lds_fft)
        for (int r = 0; r < R; r++) {
          for (int k = 0; k < Ns; k++) {
            const long double ang = -2.0L * 3.141592653589793238462643383279502884L * r * k / ((long double)Ns * R);
            t.push_back((double)cosl(ang));
            t.push_back((double)sinl(ang));
          }
        }
        Ns *= R;
      }
    }
    g.conv_tw.resize(n);  // n/2 complex
    for (int t = 0; t < n / 2; t++) {
      long double ang = -2.0L * 3.141592653589793238462643383279502884L * t / n;
      g.conv_tw[2 * t] = (double)cosl(ang);
      g.conv_tw[2 * t + 1] = (double)sinl(ang);
    }
  }
}

static int upload(DevBuf &d, const void *src, size_t bytes) {
  int e = d.ensure(bytes);
  if (e) return e;
  if (bytes) e = (int)hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice);
  return e;
}

template <class T>
static int upload_vec(DevBuf &d, const std::vector<T> &v) {
  return upload(d, v.data(), v.size() * sizeof(T));
}

int ensure_device(std::string &err) {
  Global &g = G();
  ensure_tables_host();
  if (!g.rel_loaded) {  // every model needs the relline table
    err = g.rel_err;
    return RX_ERR_TABLE;
  }
  if (!g.device_ready) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= g.device) {
      err = "no HIP device available";
      return RX_ERR_DEVICE;
    }
    e = hipSetDevice(g.device);
    if (e != hipSuccess) {
      err = "hipSetDevice failed";
      return RX_ERR_DEVICE;
    }
    int n_cu = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, g.device) == hipSuccess && n_cu > 0)
      g.n_cu = n_cu;
    // the device address of the allocation-free entry point's work array
    const int st = noalloc_setup(g.device, err);
    if (st != RX_OK) return st;
    g.device_ready = true;
  } else {
    (void)hipSetDevice(g.device);
  }
  int e = 0;
  const bool lp_dirty = g.lp_dirty && g.lp_loaded;
  if (g.rel_dirty || lp_dirty || g.conv_dirty || g.refl_dirty) {
    // resident buffers are about to be overwritten or reallocated: no kernel
    // of any stream may still read them
    if (hipDeviceSynchronize() != hipSuccess) {
      err = "device synchronisation failed";
      return RX_ERR_DEVICE;
    }
  }
  if (g.rel_dirty) {
    const RelTableHost &t = g.rel;
    if (t.n_r > MAX_NR_TAB || t.n_g > MAX_NG) {
      err = "relline table dims exceed kernel limits";
      return RX_ERR_ARG;
    }
    e |= upload_vec(g.t_a, t.a) | upload_vec(g.t_mu0, t.mu0) | upload_vec(g.t_r, t.r) |
         upload_vec(g.t_gmin, t.gmin) | upload_vec(g.t_gmax, t.gmax) | upload_vec(g.t_trff1, t.trff1) |
         upload_vec(g.t_trff2, t.trff2) | upload_vec(g.t_cosne1, t.cosne1) | upload_vec(g.t_cosne2, t.cosne2);
    g.rel_dirty = false;
  }
  if (lp_dirty) {
    const LpTableHost &t = g.lp;
    if (t.n_r > MAX_LP_R) {
      err = "lamp-post table dims exceed kernel limits";
      return RX_ERR_ARG;
    }
    e |= upload_vec(g.l_a, t.a) | upload_vec(g.l_h, t.h) | upload_vec(g.l_rad, t.rad) |
         upload_vec(g.l_int, t.intens) | upload_vec(g.l_del, t.del) | upload_vec(g.l_inc, t.del_inc);
    g.lp_dirty = false;
  }
  if (g.conv_dirty) {
    e |= upload_vec(g.c_coarse, g.coarse_ener) | upload_vec(g.c_coarse_ln, g.coarse_ln);
    e |= upload_vec(g.c_ln_emid, g.conv_ln_emid);
    e |= upload_vec(g.c_ener, g.conv_ener) | upload_vec(g.c_cfac, g.conv_cfac) | upload_vec(g.c_tw, g.conv_tw);
    e |= upload_vec(g.c_cfinv, g.conv_cfinv) | upload_vec(g.c_inv_emid, g.conv_inv_emid);
    e |= upload_vec(g.c_twp_fwd, g.conv_twp_fwd) | upload_vec(g.c_twp_inv, g.conv_twp_inv);
    g.conv_dirty = false;
  }
  if (g.refl_dirty) {
    e |= upload_vec(g.d_refl_sets, g.refl_sets);
    g.refl_dirty = false;
  }
  if (e) {
    err = "device upload failed";
    return RX_ERR_DEVICE;
  }
  return RX_OK;
}

void ensure_xilltable_host(int prim) {
  XillSlot &x = G().xill[prim];
  if (x.loaded) return;
  if (!getenv("RELXILL_TABLE_PATH")) {
    // synthetic tables: Ecut 5-1000 keV, kTe 1-400 keV on the fourth axis
    x.t = prim == PRIM_NTHCOMP ? make_synthetic_xilltable(6, 13, 5, 8, 6, 10, 2999, 1.0, 400.0)
          : prim == PRIM_BBODY ? make_synthetic_xilltable_ns(7, 12, 5, 8, 5, 10, 2999)
                               : make_synthetic_xilltable(5, 13, 5, 8, 6, 10, 2999);
    x.err.clear();
    x.loaded = true;
  } else {
    // the reference's table names (src/common.h:163, :166, :167)
    const char *name = prim == PRIM_NTHCOMP ? "xillverCp_v3.4.fits"
                       : prim == PRIM_BBODY ? "xillverNS-2.fits"
                                            : "xillver-a-Ec5.fits";
    x.loaded = load_from_path(x.t, name, nullptr, load_xilltable_fits, load_xilltable_fits, x.err);
  }
  if (x.loaded) x.dirty = true;
}

int ensure_xilltable_device(int prim, std::string &err) {
  XillSlot &x = G().xill[prim];
  ensure_xilltable_host(prim);
  const char *what = prim == PRIM_NTHCOMP ? "xillver table (nthcomp): "
                     : prim == PRIM_BBODY ? "xillver table (blackbody): "
                                          : "xillver table: ";
  if (!x.loaded) {
    err = what + x.err;
    return RX_ERR_TABLE;
  }
  if (!x.dirty) return RX_OK;
  const XillTableHost &t = x.t;
  if (t.n_incl() > MAX_INCL) {
    err = "xillver table has more inclinations than the kernels support";
    return RX_ERR_ARG;
  }
  if (hipDeviceSynchronize() != hipSuccess) {
    err = "device synchronisation failed";
    return RX_ERR_DEVICE;
  }
  // interp_xill_table's energy grid (src/xilltable.c:1104-1108)
  x.ener.resize(t.n_ener + 1);
  for (int i = 0; i < t.n_ener; i++) x.ener[i] = t.elo[i];
  x.ener[t.n_ener] = t.ehi[t.n_ener - 1];
  int e = upload_vec(x.spec, xilltable_renormed(t)) | upload_vec(x.d_ener, x.ener);
  // the inclination axis (returning-radiation corrections of relxilllp)
  e |= upload_vec(x.incl, t.val[t.n_par - 1]);
  // the logXi axis (the zones of an ionisation gradient bracket their own)
  const int k_lxi = t.dim_of(XP_LXI);
  if (k_lxi >= 0) e |= upload_vec(x.lxi, t.val[k_lxi]);
  // the Ecut / kTe axis (the lamp post's zones bracket their own value on the
  // device) and the axis ensure_ecut_within_boundarys reads
  const int k_ect = t.dim_of(XP_ECT);
  if (k_ect >= 0) {
    e |= upload_vec(x.ect, t.val[k_ect]);
    const int q = t.pid[k_ect] < t.n_par ? t.pid[k_ect] : k_ect;
    e |= upload_vec(x.q, t.val[q]);
  }
  if (e) {
    err = "xillver table upload failed";
    return RX_ERR_DEVICE;
  }
  x.dirty = false;
  return RX_OK;
}

void ensure_rradtable_host() {
  Global &g = G();
  if (g.rrad_loaded) return;
  if (!getenv("RELXILL_TABLE_PATH")) {
    g.rrad = make_synthetic_rradtable(7, 21, 50, 20);
    g.rrad_err.clear();
    g.rrad_loaded = true;
  } else {
    // the reference's table name (src/common.h:128)
    g.rrad_loaded = load_from_path(g.rrad, "table_returnRad_v20220301.fits", nullptr, load_rradtable_fits,
                                   load_rradtable_fits, g.rrad_err);
  }
  if (g.rrad_loaded) g.rrad_dirty = true;
}

int ensure_rradtable_device(std::string &err) {
  Global &g = G();
  ensure_rradtable_host();
  if (!g.rrad_loaded) {
    err = "returning-radiation table: " + g.rrad_err;
    return RX_ERR_TABLE;
  }
  if (!g.rrad_dirty) return RX_OK;
  if (hipDeviceSynchronize() != hipSuccess) {
    err = "device synchronisation failed";
    return RX_ERR_DEVICE;
  }
  const RradTableHost &t = g.rrad;
  int e = upload_vec(g.r_rlo, t.rlo) | upload_vec(g.r_rhi, t.rhi) | upload_vec(g.r_tf, t.tf_r) |
          upload_vec(g.r_gmin, t.gmin) | upload_vec(g.r_gmax, t.gmax) | upload_vec(g.r_fg, t.frac_g);
  if (e) {
    err = "returning-radiation table upload failed";
    return RX_ERR_DEVICE;
  }
  g.rrad_dirty = false;
  return RX_OK;
}

int rrad_zones(const RradTableHost &t, ReqParams &p, int nr_fine) {
  // the fine grid's ends (determine_rlo_rhi, src/Relreturn_Corona.cpp:201-213)
  const double rin = fine_radius(p.rin, p.rout, nr_fine - 1, nr_fine);
  const double rout = fine_radius(p.rin, p.rout, 0, nr_fine);
  int k = bsearch_asc(t.spin.data(), t.n_spin, p.a);
  if (k < 0) k = 0;
  if (t.spin[k] < p.a) k++;
  if (k >= t.n_spin) return 35;
  const int nt = t.n_rad;
  const double *rlo = &t.rlo[(size_t)k * nt], *rhi = &t.rhi[(size_t)k * nt];
  int klo = bsearch_asc(rlo, nt, rin);
  int khi = bsearch_asc(rhi, nt, rout);
  if (fabs(rout - rhi[nt - 1]) < 1e-6) khi = nt - 1; else khi++;
  if (fabs(rin - rlo[0]) < 1e-6) klo = 0;
  const int n = (khi + 1) - klo;
  if (n < 2 || n > nt || n > MAX_RR || klo < 0) return 36;
  if (!(rin >= rlo[klo] - 1e-4)) return 37;
  p.rr_k = k;
  p.rr_klo = klo;
  p.rr_n = n;
  return 0;
}

void fill_dev_views(BatchDev &bd) {
  Global &g = G();
  bd.rt.n_a = g.rel.n_a;
  bd.rt.n_mu0 = g.rel.n_mu0;
  bd.rt.n_r = g.rel.n_r;
  bd.rt.n_g = g.rel.n_g;
  bd.rt.a = g.t_a.as<float>();
  bd.rt.mu0 = g.t_mu0.as<float>();
  bd.rt.r = g.t_r.as<float>();
  bd.rt.gmin = g.t_gmin.as<float>();
  bd.rt.gmax = g.t_gmax.as<float>();
  bd.rt.trff1 = g.t_trff1.as<float>();
  bd.rt.trff2 = g.t_trff2.as<float>();
  bd.rt.cosne1 = g.t_cosne1.as<float>();
  bd.rt.cosne2 = g.t_cosne2.as<float>();
  bd.lp.n_a = g.lp.n_a;
  bd.lp.n_h = g.lp.n_h;
  bd.lp.n_r = g.lp.n_r;
  bd.lp.a = g.l_a.as<float>();
  bd.lp.h = g.l_h.as<float>();
  bd.lp.rad = g.l_rad.as<float>();
  bd.lp.intens = g.l_int.as<float>();
  bd.lp.del = g.l_del.as<float>();
  bd.lp.del_inc = g.l_inc.as<float>();
  bd.cg.n = g.conv_n;
  bd.cg.i1kev = g.conv_i1kev;
  bd.cg.band_lo = g.conv_band_lo;
  bd.cg.band_hi = g.conv_band_hi;
  bd.cg.ener = g.c_ener.as<double>();
  bd.cg.cfac = g.c_cfac.as<double>();
  bd.cg.cfinv = g.c_cfinv.as<double>();
  bd.cg.inv_emid = g.c_inv_emid.as<double>();
  bd.cg.twp_fwd = g.c_twp_fwd.as<double>();
  bd.cg.twp_inv = g.c_twp_inv.as<double>();
  bd.cg.tw = g.c_tw.as<double>();
  bd.cg.coarse = g.c_coarse.as<double>();
  bd.cg.coarse_ln = g.c_coarse_ln.as<double>();
  bd.cg.ln_emid = g.c_ln_emid.as<double>();
  bd.refl_sets = g.d_refl_sets.as<DevReflSet>();
  for (int pr = 0; pr < N_PRIM; pr++) {
    const XillSlot &x = g.xill[pr];
    const XillTableHost &t = x.t;
    DevXillTable &xt = bd.xt[pr];
    const bool ok = x.loaded && t.n_par > 0;
    xt.n_ener = ok ? t.n_ener : 0;
    xt.n_incl = ok ? t.n_incl() : 0;
    // six parameters: the first is interpolated outside the four inner ones
    const int off = ok ? t.n_par - 5 : 0;
    for (int k = 0; k < 4; k++) xt.n_val[k] = ok ? t.n_val[k + off] : 0;
    xt.n_out = (ok && off) ? t.n_val[0] : 1;
    xt.outer_rows = 0;
    if (ok && off) {
      xt.outer_rows = 1;
      for (int k = 1; k < t.n_par; k++) xt.outer_rows *= (unsigned)t.n_val[k];
    }
    xt.spec = x.spec.as<float>();
    xt.ener = x.d_ener.as<double>();
    // the Ecut axis among the inner four (never the outer one: validate_xilltable)
    const int k_ect = ok ? t.dim_of(XP_ECT) : -1;
    xt.ect_dim = k_ect >= 0 ? k_ect - off : -1;
    xt.ect_q = -1;
    xt.n_ect = xt.n_q = 0;
    if (k_ect >= 0) {
      const int q = t.pid[k_ect] < t.n_par ? t.pid[k_ect] : k_ect;  // table axes (x.ect, x.q)
      xt.ect_q = q;
      xt.n_ect = t.n_val[k_ect];
      xt.n_q = t.n_val[q];
    }
    xt.ect_vals = x.ect.as<float>();
    xt.incl_vals = x.incl.as<float>();
    xt.q_vals = x.q.as<float>();
    const int k_lxi = ok ? t.dim_of(XP_LXI) : -1;
    xt.lxi_dim = (k_lxi >= off) ? k_lxi - off : -1;  // (a first axis of six is bracketed on the host)
    xt.n_lxi = k_lxi >= 0 ? t.n_val[k_lxi] : 0;
    xt.lxi_vals = x.lxi.as<float>();
  }
  bd.rr.n_spin = g.rrad.n_spin;
  bd.rr.n_rad = g.rrad.n_rad;
  bd.rr.n_g = g.rrad.n_g;
  bd.rr.rlo = g.r_rlo.as<double>();
  bd.rr.rhi = g.r_rhi.as<double>();
  bd.rr.tf_r = g.r_tf.as<double>();
  bd.rr.gmin = g.r_gmin.as<double>();
  bd.rr.gmax = g.r_gmax.as<double>();
  bd.rr.frac_g = g.r_fg.as<double>();
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------

Batch::~Batch() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

static bool is_relxill(int model) { return model == MOD_RELXILL || model == MOD_RELXILLLP; }

// the inclination parameter (degrees) of a relxill model
static double incl_deg_of(int model, const double *par) { return par[model == MOD_RELXILLLP ? 3 : 4]; }

static bool is_conv(int model) {
  return model == MOD_RELCONV || model == MOD_RELCONVLP || model == MOD_RELXILL_CONV || is_relxill(model);
}

// The nthcomp-primary models run as relxill / relxilllp requests
// (convertModelType maps relxillCp to MOD_TYPE_RELXILL and relxilllpCp to
// MOD_TYPE_RELXILLLP, src/ModelDefinition.cpp:38-50; their prim_type is
// PRIM_SPEC_NTHCOMP, :88): their parameters in those models' order
// (src/modelfiles/lmodel_relxill_public.dat:96-129) with kTe in Ecut's place,
// and logN.  Returns the model to run, or 0 for other models.
struct CpRequest {
  int model = 0;
  double par[14] = {};
  double dens = 15.0;
  double ktbb = 0.0;
  int prim = PRIM_NTHCOMP;
  int status = RX_OK;  // RX_ERR_UNSUPPORTED: an ionisation gradient that is not built was asked for
  int ion_grad = 0;     // relxilllpCp's iongrad_type (1: the power law)
  double ion_index = 0.0;
};

static CpRequest remap_cp(int model, const double *par) {
  CpRequest c;
  if (model == MOD_RELXILLCP) {
    // Incl a Rin Rout Rbr Index1 Index2 z gamma logxi logN Afe kTe refl_frac
    static const int k[13] = {5, 6, 4, 1, 0, 2, 3, 7, 8, 9, 11, 12, 13};
    for (int i = 0; i < 13; i++) c.par[i] = par[k[i]];
    c.dens = par[10];
    c.model = MOD_RELXILL;
  } else if (model == MOD_RELXILLLPCP) {
    // Incl a Rin Rout h beta gamma logxi logN Afe kTe refl_frac z iongrad_index
    // iongrad_type switch_returnrad switch_reflfrac_boost
    static const int k[14] = {4, 5, 1, 0, 2, 3, 12, 6, 7, 9, 10, 11, 15, 16};
    for (int i = 0; i < 14; i++) c.par[i] = par[k[i]];
    c.dens = par[8];
    c.model = MOD_RELXILLLP;
    // iongrad_type (src/IonGradient.cpp:340-395): 0 constant ionisation, 1 the
    // power law in radius (calc_ion_grad_pl, :174-182) with iongrad_index;
    // the alpha-disc gradient (2) is not built
    const long ig = lround(par[14]);
    if (ig == 1) {
      c.ion_grad = 1;
      c.ion_index = par[13];
    } else if (ig != 0) {
      c.status = RX_ERR_UNSUPPORTED;
    }
  } else if (model == MOD_XILLVERCP) {
    // xillverCp (lmodel_relxill_public.dat:86-94): gamma Afe kTe logxi logN z
    // Incl refl_frac -- xillver's order with logN taken out
    static const int k[7] = {0, 1, 2, 3, 5, 6, 7};
    for (int i = 0; i < 7; i++) c.par[i] = par[k[i]];
    c.dens = par[4];
    c.model = MOD_XILLVER;
  } else if (model == MOD_XILLVERNS) {
    // xillverNS (:131-138): kTbb Afe logN logxi z Incl refl_frac; Gamma 0,
    // Ecut 300 as for relxillNS
    const double v[7] = {0.0, par[1], 300.0, par[3], par[4], par[5], par[6]};
    for (int i = 0; i < 7; i++) c.par[i] = v[i];
    c.ktbb = par[0];
    c.dens = par[2];
    c.model = MOD_XILLVER;
    c.prim = PRIM_BBODY;
  } else if (model == MOD_RELXILLNS) {
    // relxillNS (lmodel_relxill_public.dat:140-153): Index1 Index2 Rbr a Incl
    // Rin Rout z kTbb logxi Afe logN refl_frac -- relxill's order with kTbb in
    // Gamma's slot and logN in Ecut's; get_xill_params leaves Gamma 0 and Ecut
    // at its default 300 (src/ModelDefinition.cpp:357-376)
    for (int i = 0; i < 13; i++) c.par[i] = par[i];
    c.ktbb = par[8];
    c.dens = par[11];
    c.par[8] = 0.0;
    c.par[11] = 300.0;
    c.model = MOD_RELXILL;
    c.prim = PRIM_BBODY;
  }
  return c;
}

int batch_build(Batch &b, const rx_request *req_in, int n_req, std::string &err) {
  Global &g = G();
  if (n_req <= 0 || !req_in) {
    err = "empty batch";
    return RX_ERR_ARG;
  }
  // nthcomp-primary requests in their relxill form
  std::vector<rx_request> req_local;
  std::vector<CpRequest> cp;
  std::vector<int> prim(n_req, PRIM_CUTOFFPL);
  const rx_request *req = req_in;
  for (int i = 0; i < n_req; i++) {
    const int m = req_in[i].model;
    if ((m == MOD_RELXILLCP || m == MOD_RELXILLLPCP || m == MOD_RELXILLNS || m == MOD_XILLVERCP ||
         m == MOD_XILLVERNS) &&
        req_in[i].parameter) {
      if (cp.empty()) {
        req_local.assign(req_in, req_in + n_req);
        cp.resize(n_req);
        req = req_local.data();
      }
      cp[i] = remap_cp(m, req_in[i].parameter);
      req_local[i].model = cp[i].model;
      req_local[i].parameter = cp[i].par;
      prim[i] = cp[i].prim;
    }
  }
  const bool phys = phys_norm_env();
  b.n_req = n_req;
  b.h_req.assign(n_req, ReqParams());
  b.h_status.assign(n_req, 0);
  b.h_conv.clear();
  b.h_xill.clear();
  b.h_prim.clear();
  b.h_rrc.clear();
  b.h_nth.clear();
  b.n_nth_sol = 0;
  std::vector<XillWork> xill_multi;
  for (int pr = 0; pr < N_PRIM; pr++) {
    for (int i = 0; i < n_req; i++) {
      if ((is_relxill(req[i].model) || req[i].model == MOD_XILLVER) && prim[i] == pr) {
        int st = ensure_xilltable_device(pr, err);
        if (st != RX_OK) return st;
        break;
      }
    }
  }
  for (int i = 0; i < n_req; i++) {
    const int m = req[i].model;
    if ((m == MOD_RELLINELP || m == MOD_RELCONVLP || m == MOD_RELXILLLP) && !g.lp_loaded) {
      err = "lamp-post table: " + g.lp_err;
      return RX_ERR_TABLE;
    }
  }
  for (int i = 0; i < n_req; i++) {
    const int m = req[i].model;
    const int k_rr = (m == MOD_RELLINELP) ? 9 : (m == MOD_RELCONVLP) ? 8 : (m == MOD_RELXILLLP) ? 12 : -1;
    if (k_rr >= 0 && req[i].parameter && lround(req[i].parameter[k_rr]) != 0) {
      int st = ensure_rradtable_device(err);
      if (st != RX_OK) return st;
      break;
    }
  }
  int64_t o_ener = 0, o_flux = 0, o_pgrid = 0, o_prof = 0, o_xspec = 0, o_dist = 0, o_zone = 0, o_rrc = 0;
  int64_t o_node = 0;
  const int64_t node_stride = 2 * (int64_t)g.rel.n_r * g.rel.n_g * 2;
  for (int i = 0; i < n_req; i++) {
    const rx_request &r = req[i];
    if (r.n_flux <= 0 || !r.energy || !r.parameter) {
      err = "request with empty grid or missing parameters";
      return RX_ERR_ARG;
    }
    if (r.n_flux > (int)RAD_BIN_MASK) {  // bins are 24-bit fields of the radius records
      err = "energy grid of more than 2^24 - 1 bins";
      return RX_ERR_ARG;
    }
    ReqParams &p = b.h_req[i];
    std::string perr;
    int st = parse_request(r.model, r.parameter, phys, p, perr, prim[i]);
    if (st == RX_ERR_ARG) {
      err = perr;
      return st;
    }
    if (st == RX_OK && prim[i] != PRIM_CUTOFFPL) st = cp[i].status;
    if (st == RX_OK && prim[i] != PRIM_CUTOFFPL && cp[i].ion_grad == 1) {
      // an ionisation gradient: N_ZONES_IONGRAD zones (get_num_zones,
      // src/relutility.c:517-527) whose logXi the device sets from the input
      // one (zone_record)
      p.n_zones = MAX_ZONES;
      rzone_grid(p.rin, p.rout, p.n_zones, p.height, p.zgrid);
      p.ion_grad = 1;
      p.ion_index = cp[i].ion_index;
      p.xlxi0 = cp[i].par[8];  // relxilllp's logxi slot, before the 0-4.7 clamp
    }
    p.n_in = r.n_flux;
    p.ener_off = o_ener;
    o_ener += r.n_flux + 1;
    p.flux_off = o_flux;
    o_flux += r.n_flux;
    p.refl_set = r.refl_set;
    if (p.n_zones < 1) p.n_zones = 1;
    if (r.model == MOD_XILLVER) {
      // the xillver models: the table at the parameters with the inclination
      // interpolated as a fifth axis (k_xill over a two-bin angle
      // distribution), rebinned with the primary (k_primary)
      p.n_e = 0;
      p.pgrid_off = -1;
      p.prof_off = -1;
      const XillTableHost &xt = g.xill[p.prim].t;
      const double dens = p.prim != PRIM_CUTOFFPL ? cp[i].dens : 15.0;
      p.xktbb = p.prim == PRIM_BBODY ? cp[i].ktbb : 0.0;
      xill_corner(xt, p.xgam, r.parameter[1], r.parameter[3], p.xect_src, dens, p.xktbb, r.parameter[5], p.xind,
                  p.xfac, true, &p.xinc_ind, &p.xinc_fac, &p.xo_ind, &p.xo_fac);
      p.xdens_div = (xt.dim_of(XP_DNS) < 0 && std::fabs(dens - 15) > 1e-6) ? std::pow(10, dens - 15) : 0.0;
      p.n_incl = xt.n_incl();
      p.xspec_off = o_xspec;
      o_xspec += (int64_t)xt.n_ener;
      p.zone_off = o_zone;
      o_zone += 1;
      b.h_xill.push_back(XillWork{i, 0});
      b.h_prim.push_back(i);
      if (p.prim == PRIM_NTHCOMP) {
        p.n_nth = 1;
        p.nth_off = (int64_t)b.h_nth.size();
        p.nth_sol = b.n_nth_sol++;
        b.h_nth.push_back(XillWork{i, 0});
      }
    } else if (is_conv(r.model)) {
      p.n_e = g.conv_n;
      p.pgrid_off = -1;
      p.prof_off = o_prof;
      o_prof += (int64_t)g.conv_n * p.n_zones;
      for (int z = 0; z < p.n_zones; z++) b.h_conv.push_back(XillWork{i, z});
      if (is_relxill(r.model)) {
        // the zones' xillver parameters (src/IonGradient.cpp:196-209, :340-389):
        // logXi kept within 0-4.7, density 15; the Ecut axis of the lamp
        // post's zones is bracketed on the device (it depends on the zone's
        // energy shift)
        const bool lp = r.model == MOD_RELXILLLP;
        double lxi = r.parameter[lp ? 8 : 9];
        if (lxi < 0.0) lxi = 0.0;
        else if (lxi > 4.7) lxi = 4.7;
        const double afe = r.parameter[lp ? 9 : 10];
        const XillTableHost &xt = g.xill[p.prim].t;
        // logN: 15 for the cut-off power-law models (get_xill_params,
        // src/ModelDefinition.cpp:367-368), the parameter for the others
        const double dens = p.prim != PRIM_CUTOFFPL ? cp[i].dens : 15.0;
        p.xktbb = p.prim == PRIM_BBODY ? cp[i].ktbb : 0.0;
        xill_corner(xt, p.xgam, afe, lxi, p.xect_src, dens, p.xktbb, incl_deg_of(r.model, r.parameter), p.xind,
                    p.xfac, false, nullptr, nullptr, &p.xo_ind, &p.xo_fac);
        // a table without a density axis divides its rows by 10^(logN - 15)
        // (normalizeXillverSpecLogxiDensity, src/xilltable.c:488-511)
        p.xdens_div = (xt.dim_of(XP_DNS) < 0 && std::fabs(dens - 15) > 1e-6) ? std::pow(10, dens - 15) : 0.0;
        p.n_incl = xt.n_incl();
        p.xspec_off = o_xspec;
        o_xspec += (int64_t)xt.n_ener * p.n_zones;
        p.zone_off = o_zone;
        o_zone += p.n_zones;
        if (p.ion_grad) {
          // zones that differ in logXi as well as in Ecut: one k_xill1 unit each
          for (int z = 0; z < p.n_zones; z++) b.h_xill.push_back(XillWork{i, z});
        } else {
          (p.n_zones > 1 ? xill_multi : b.h_xill).push_back(XillWork{i, 0});
        }
        b.h_prim.push_back(i);
        if (p.prim == PRIM_NTHCOMP) {
          // the source's solve, then one per lamp-post zone (k_nth)
          p.n_nth = 1 + (p.emis_type == EMIS_LP ? p.n_zones : 0);
          p.nth_off = (int64_t)b.h_nth.size();
          p.nth_sol = b.n_nth_sol++;
          for (int j = 0; j < p.n_nth; j++) b.h_nth.push_back(XillWork{i, j});
        }
      } else if (st == RX_OK && r.model != MOD_RELXILL_CONV) {
        if (!r.flux_in) {
          err = "convolution request without input spectrum";
          return RX_ERR_ARG;
        }
        double sum = 0.0;  // src/LocalModel.cpp:84-86
        for (int k = 0; k < r.n_flux; k++) sum += r.flux_in[k];
        if (!(sum > 0.0)) st = RX_ERR_INPUT;
      }
      if (st == RX_OK && r.model == MOD_RELXILL_CONV) {
        if (r.refl_set < 0 || r.refl_set >= (int)g.refl_sets.size()) {
          err = "unknown reflection set";
          return RX_ERR_ARG;
        }
        p.n_incl = g.refl_sets[r.refl_set].n_incl;
      }
    } else {
      p.n_e = r.n_flux;
      p.pgrid_off = o_pgrid;
      o_pgrid += r.n_flux + 1;
    }
    if (st == RX_OK && p.return_rad != 0) st = rrad_zones(g.rrad, p, g.nr_fine);
    p.rr_corr_off = 0;
    if (st == RX_OK && p.rr_corr) {
      if (g.xill[p.prim].t.n_ener > MAX_RRC_ENER) {
        err = "xillver table has more energy bins than the returning-radiation corrections hold";
        return RX_ERR_ARG;
      }
      p.rr_corr_off = o_rrc;
      o_rrc += 2 * (int64_t)p.n_zones;
      for (int z = 0; z < p.n_zones; z++) b.h_rrc.push_back(XillWork{i, z});
    }
    const bool xill_model = r.model == MOD_XILLVER;
    p.need_cosne = (p.limb != 0 || (p.n_incl > 0 && !xill_model)) ? 1 : 0;
    p.dist_off = o_dist;
    o_dist += (int64_t)p.n_incl * p.n_zones;
    p.node_off = -1;
    if (p.n_incl > 0 && !xill_model) {
      p.node_off = o_node;
      o_node += node_stride;
    }
    p.pre_status = st;
    b.h_status[i] = st;
  }
  b.tot_ener = o_ener;
  b.tot_flux = o_flux;
  b.tot_pgrid = o_pgrid;
  b.tot_prof = o_prof;
  b.tot_xspec = o_xspec;
  b.tot_dist = o_dist;
  b.tot_node = o_node;
  b.tot_zone = o_zone;
  b.tot_rrc = o_rrc;
  b.n_xill1 = (int)b.h_xill.size();
  b.h_xill.insert(b.h_xill.end(), xill_multi.begin(), xill_multi.end());
  b.max_zones = 1;
  for (int i = 0; i < n_req; i++) b.max_zones = std::max(b.max_zones, b.h_req[i].n_zones);
  // k_conv takes one workgroup per unit, in this order: the units whose
  // caller grids are longest (their rebins cost most) go first, so that the
  // last round of workgroups is of short ones (0.39 -> 0.36 ms on 4096
  // requests of 300-8000 bins)
  std::stable_sort(b.h_conv.begin(), b.h_conv.end(), [&](const XillWork &x, const XillWork &y) {
    return b.h_req[x.req].n_in > b.h_req[y.req].n_in;
  });
  b.h_ener.resize(o_ener);
  b.h_flux_in.assign(o_flux, 0.0);
  for (int i = 0; i < n_req; i++) {
    const rx_request &r = req[i];
    const ReqParams &p = b.h_req[i];
    memcpy(&b.h_ener[p.ener_off], r.energy, sizeof(double) * (r.n_flux + 1));
    if ((is_relxill(r.model) || r.model == MOD_XILLVER) && p.z > 0) {
      // relxill evaluates on the redshifted grid (shift_energy_grid_redshift,
      // src/XspecSpectrum.h:70-76)
      for (int k = 0; k <= r.n_flux; k++) b.h_ener[p.ener_off + k] *= (1 + p.z);
    }
    if (r.flux_in && is_conv(r.model) && !is_relxill(r.model))
      memcpy(&b.h_flux_in[p.flux_off], r.flux_in, sizeof(double) * r.n_flux);
  }
  b.any_cosne = false;
  b.any_limb = false;
  int64_t bins = 0;
  for (int i = 0; i < n_req; i++) {
    b.any_cosne |= b.h_req[i].need_cosne != 0;
    b.any_limb |= b.h_req[i].limb != 0;
    bins += b.h_req[i].n_e;
  }
  // shape class for the autotuner: log2 of the total profile bins, conv flag
  int lg = 0;
  while ((1LL << lg) < bins) lg++;
  b.shape_class = lg * 2 + (b.h_conv.empty() ? 0 : 1);
  batch_plan(b, b.seg_forced, b.waves_forced);
  b.config_gen = g.config_gen;
  return batch_upload(b);
}

void batch_plan(Batch &b, int seg_bins, int waves_per_cu) {
  Global &g = G();
  if (seg_bins <= 0 || waves_per_cu <= 0) {
    for (auto &t : g.tune) {
      if (t.shape == b.shape_class) {
        if (seg_bins <= 0) seg_bins = t.seg_bins;
        if (waves_per_cu <= 0) waves_per_cu = t.waves_per_cu;
      }
    }
  }
  if (seg_bins <= 0) seg_bins = 64;
  if (waves_per_cu <= 0) waves_per_cu = 8;
  b.seg_bins = seg_bins;
  b.waves_per_cu = waves_per_cu;
  // work list bound: every (radius chunk, zone) pair of every request covering
  // all its bins; the zones are contiguous radius ranges, so a request has at
  // most n_rc + n_zones - 1 such pairs
  const int n_rc = (g.nr_fine + RAD_CHUNK - 1) / RAD_CHUNK;
  int64_t cap = 0;
  for (int i = 0; i < b.n_req; i++)
    if (b.h_status[i] == RX_OK)
      cap += (int64_t)(n_rc + b.h_req[i].n_zones - 1) * ((b.h_req[i].n_e + seg_bins - 1) / seg_bins);
  b.item_cap = (int)std::min<int64_t>(std::max<int64_t>(cap, 1), INT32_MAX);
}

int batch_upload(Batch &b) {
  Global &g = G();
  int e = 0;
  if (!b.stream) {
    e |= (int)hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking);
    for (auto &x : b.ev) e |= (int)hipEventCreate(&x);
  }
  const size_t nreq = b.n_req;
  e |= b.d_req.ensure(sizeof(ReqParams) * nreq);
  e |= b.d_ener.ensure(sizeof(double) * b.tot_ener);
  e |= b.d_flux_in.ensure(sizeof(double) * b.tot_flux);
  e |= b.d_out.ensure(sizeof(double) * b.tot_flux);
  e |= b.d_pgrid.ensure(sizeof(double) * std::max<int64_t>(1, b.tot_pgrid));
  e |= b.d_node.ensure(sizeof(double) * std::max<int64_t>(1, b.tot_node));
  e |= b.d_rad.ensure(sizeof(RadRec) * nreq * g.nr_fine);
  e |= b.d_rnode.ensure(sizeof(double) * RAD_NODE_ARRAYS * nreq * g.rel.n_r);
  e |= b.d_rnorm.ensure(sizeof(ReqNorm) * nreq);
  e |= b.d_prof.ensure(sizeof(double) * std::max<int64_t>(1, b.tot_prof));
  e |= b.d_dist.ensure(sizeof(double) * std::max<int64_t>(1, b.tot_dist));
  e |= b.d_zone.ensure(sizeof(ZoneRec) * std::max<int64_t>(1, b.tot_zone));
  e |= b.d_refl.ensure(sizeof(LpRefl) * nreq);
  e |= b.d_status.ensure(sizeof(int) * nreq);
  e |= b.d_conv.ensure(sizeof(XillWork) * std::max<size_t>(1, b.h_conv.size()));
  e |= b.d_xill.ensure(sizeof(XillWork) * std::max<size_t>(1, b.h_xill.size()));
  e |= b.d_prim.ensure(sizeof(int) * std::max<size_t>(1, b.h_prim.size()));
  e |= b.d_rrc.ensure(sizeof(XillWork) * std::max<size_t>(1, b.h_rrc.size()));
  e |= b.d_rrcorr.ensure(sizeof(double) * std::max<int64_t>(1, b.tot_rrc));
  e |= b.d_nthw.ensure(sizeof(XillWork) * std::max<size_t>(1, b.h_nth.size()));
  e |= b.d_nth.ensure(sizeof(NthRes) * std::max<size_t>(1, b.h_nth.size()));
  e |= b.d_nthx.ensure(sizeof(double) * (NTH_MAX_POINTS + 1) * std::max(1, b.n_nth_sol));
  e |= b.d_nths.ensure(sizeof(double) * NTH_MAX_POINTS * std::max(1, b.n_nth_sol));
  e |= b.d_nthn.ensure(sizeof(int) * std::max(1, b.n_nth_sol));
  e |= b.d_xspec.ensure(sizeof(double) * std::max<int64_t>(1, b.tot_xspec));
  e |= b.d_items.ensure(sizeof(WorkItem) * (size_t)b.item_cap);
  e |= b.d_ctr.ensure(sizeof(int) * N_COUNTERS);
  e |= b.d_reach.ensure(sizeof(BinRange) * nreq);
  // FP64 re-check queue: a few percent of the bins (overflow redoes the batch in FP64)
  int64_t bins = 0;
  for (int i = 0; i < b.n_req; i++) bins += b.h_req[i].n_e;
  b.fix_cap = (int)std::min<int64_t>(1 << 22, std::max<int64_t>(1 << 16, bins / 8));
  e |= b.d_fix.ensure(sizeof(FixEntry) * (size_t)b.fix_cap);
  if (e) return RX_ERR_NOMEM;
  e |= (int)hipMemcpyAsync(b.d_req.p, b.h_req.data(), sizeof(ReqParams) * nreq, hipMemcpyHostToDevice, b.stream);
  e |= (int)hipMemcpyAsync(b.d_ener.p, b.h_ener.data(), sizeof(double) * b.tot_ener, hipMemcpyHostToDevice, b.stream);
  e |= (int)hipMemcpyAsync(b.d_flux_in.p, b.h_flux_in.data(), sizeof(double) * b.tot_flux, hipMemcpyHostToDevice, b.stream);
  e |= (int)hipMemcpyAsync(b.d_status.p, b.h_status.data(), sizeof(int) * nreq, hipMemcpyHostToDevice, b.stream);
  if (!b.h_conv.empty())
    e |= (int)hipMemcpyAsync(b.d_conv.p, b.h_conv.data(), sizeof(XillWork) * b.h_conv.size(), hipMemcpyHostToDevice,
                             b.stream);
  if (!b.h_rrc.empty())
    e |= (int)hipMemcpyAsync(b.d_rrc.p, b.h_rrc.data(), sizeof(XillWork) * b.h_rrc.size(), hipMemcpyHostToDevice,
                             b.stream);
  if (!b.h_prim.empty())
    e |= (int)hipMemcpyAsync(b.d_prim.p, b.h_prim.data(), sizeof(int) * b.h_prim.size(), hipMemcpyHostToDevice, b.stream);
  if (!b.h_xill.empty())
    e |= (int)hipMemcpyAsync(b.d_xill.p, b.h_xill.data(), sizeof(XillWork) * b.h_xill.size(), hipMemcpyHostToDevice,
                             b.stream);
  if (!b.h_nth.empty())
    e |= (int)hipMemcpyAsync(b.d_nthw.p, b.h_nth.data(), sizeof(XillWork) * b.h_nth.size(), hipMemcpyHostToDevice,
                             b.stream);
  e |= (int)hipStreamSynchronize(b.stream);
  return e ? RX_ERR_DEVICE : RX_OK;
}

static BatchDev dev_view(Batch &b) {
  Global &g = G();
  BatchDev bd;
  memset(&bd, 0, sizeof(bd));
  bd.n_req = b.n_req;
  bd.nr_fine = g.nr_fine;
  bd.seg_bins = b.seg_bins;
  bd.node_stride = 2 * g.rel.n_r * g.rel.n_g * 2;
  bd.max_zones = b.max_zones;
  bd.req = b.d_req.as<ReqParams>();
  bd.ener_in = b.d_ener.as<double>();
  bd.flux_in = b.d_flux_in.as<double>();
  bd.out = b.d_out.as<double>();
  bd.pgrid = b.d_pgrid.as<double>();
  bd.node = b.d_node.as<double>();
  bd.rad = b.d_rad.as<RadRec>();
  bd.rnode = b.d_rnode.as<double>();
  bd.rnorm = b.d_rnorm.as<ReqNorm>();
  bd.prof = b.d_prof.as<double>();
  bd.dist = b.d_dist.as<double>();
  bd.refl = b.d_refl.as<LpRefl>();
  bd.status = b.d_status.as<int>();
  bd.items = b.d_items.as<WorkItem>();
  bd.item_cap = b.item_cap;
  bd.ctr = b.d_ctr.as<int>();
  bd.reach = b.d_reach.as<BinRange>();
  bd.fix = b.d_fix.as<FixEntry>();
  bd.fix_cap = b.fix_cap;
  bd.mode = b.last_mode;
  bd.guard = g.guard;
  fill_dev_views(bd);
  bd.xspec = b.d_xspec.as<double>();
  bd.zone = b.d_zone.as<ZoneRec>();
  bd.rrcorr = b.d_rrcorr.as<double>();
  bd.nth = b.d_nth.as<NthRes>();
  bd.nth_x = b.d_nthx.as<double>();
  bd.nth_s = b.d_nths.as<double>();
  bd.nth_n = b.d_nthn.as<int>();
  return bd;
}

// The run's kernels in stream order (events around them when profiling).
static int launch_sequence(Batch &b, const RunKey &k, hipStream_t s) {
  const BatchDev &bd = k.bd;
  const bool prof = k.prof != 0;
  int e = 0;
  if (prof) e |= (int)hipEventRecord(b.ev[0], s);
  e |= launch_nth(bd, static_cast<const XillWork *>(k.d_nth), k.n_nth, s);
  e |= launch_rrad_corr(bd, static_cast<const XillWork *>(k.d_rrc), k.n_rrc, s);
  e |= launch_setup(bd, s);
  if (prof) e |= (int)hipEventRecord(b.ev[1], s);
  e |= launch_integrate(bd, k.n_blocks, k.cos_rows != 0, s);
  if (prof) e |= (int)hipEventRecord(b.ev[2], s);
  e |= launch_fixup(bd, k.cos_rows != 0, s);
  if (prof) e |= (int)hipEventRecord(b.ev[3], s);
  e |= launch_finalize(bd, s);
  if (prof) e |= (int)hipEventRecord(b.ev[4], s);
  e |= launch_xill(bd, static_cast<const XillWork *>(k.d_xill), k.n_xill1, k.n_xill - k.n_xill1, s);
  if (prof) e |= (int)hipEventRecord(b.ev[5], s);
  e |= launch_conv(bd, static_cast<const XillWork *>(k.d_conv), k.n_conv, s);
  if (prof) e |= (int)hipEventRecord(b.ev[6], s);
  e |= launch_primary(bd, static_cast<const int *>(k.d_prim), k.n_prim, s);
  if (prof) e |= (int)hipEventRecord(b.ev[7], s);
  return e;
}

static void drop_graph(Batch &b) {
  if (b.gexec) (void)hipGraphExecDestroy(b.gexec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  b.gexec = nullptr;
  b.graph = nullptr;
}

// Captures the launch sequence of k into the batch's graph (0 on success).
// The stream is back to normal whatever happens.
static int capture_graph(Batch &b, const RunKey &k, hipStream_t s) {
  drop_graph(b);
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  const int e = launch_sequence(b, k, s);
  hipGraph_t gr = nullptr;
  const hipError_t ec = hipStreamEndCapture(s, &gr);
  if (e || ec != hipSuccess || !gr) {
    if (gr) (void)hipGraphDestroy(gr);
    (void)hipGetLastError();
    return -1;
  }
  hipGraphExec_t ex = nullptr;
  if (hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) != hipSuccess || !ex) {
    (void)hipGraphDestroy(gr);
    (void)hipGetLastError();
    return -1;
  }
  b.graph = gr;
  b.gexec = ex;
  b.graph_key = k;
  return 0;
}

int batch_run(Batch &b, int precision, std::string &err) {
  Global &g = G();
  if (b.config_gen != g.config_gen) {
    err = "tables, convolution grid, fine-grid size or reflection sets changed since the batch was built";
    return RX_ERR_ARG;
  }
  b.last_mode = (precision == RX_FP32) ? 1 : 0;
  RunKey k;
  memset(&k, 0, sizeof(k));
  k.bd = dev_view(b);
  // persistent wavefronts: CUs x waves per CU, but no more than the list can
  // hold items (a small batch -- the server's -- would start thousands that
  // find the list empty)
  k.n_blocks = (int)std::min<int64_t>((int64_t)g.n_cu * b.waves_per_cu, std::max(1, b.item_cap));
  k.n_xill1 = b.n_xill1;
  k.n_xill = (int)b.h_xill.size();
  k.n_conv = (int)b.h_conv.size();
  k.n_prim = (int)b.h_prim.size();
  k.d_xill = b.d_xill.p;
  k.d_conv = b.d_conv.p;
  k.d_prim = b.d_prim.p;
  k.d_rrc = b.d_rrc.p;
  k.n_rrc = (int)b.h_rrc.size();
  k.d_nth = b.d_nthw.p;
  k.n_nth = (int)b.h_nth.size();
  k.prof = g.profile ? 1 : 0;
  k.cos_rows = b.any_limb ? 1 : 0;
  hipStream_t s = b.stream;
  b.profiled = k.prof != 0;
  if (g.use_graphs && !b.graph_failed && !b.one_shot) {
    if (b.gexec && memcmp(&k, &b.graph_key, sizeof(k)) == 0) {
      if (hipGraphLaunch(b.gexec, s) == hipSuccess) {
        b.graph_replays++;
        return RX_OK;
      }
      (void)hipGetLastError();
      drop_graph(b);
      b.graph_failed = true;
    } else if (b.last_valid && memcmp(&k, &b.last_key, sizeof(k)) == 0) {
      // the second run with this configuration: capture it, then replay
      if (capture_graph(b, k, s) == 0 && hipGraphLaunch(b.gexec, s) == hipSuccess) {
        b.graph_replays++;
        return RX_OK;
      }
      (void)hipGetLastError();
      drop_graph(b);
      b.graph_failed = true;  // direct launches for this batch from now on
    }
  }
  const int e = launch_sequence(b, k, s);
  b.last_key = k;
  b.last_valid = true;
  if (e) {
    err = "kernel launch failed";
    return RX_ERR_DEVICE;
  }
  return RX_OK;
}

int batch_sync(Batch &b) {
  if (!b.stream) return RX_OK;
  if (hipStreamSynchronize(b.stream) != hipSuccess) return RX_ERR_DEVICE;
  int ctr[N_COUNTERS] = {};
  if (hipMemcpy(ctr, b.d_ctr.p, sizeof ctr, hipMemcpyDeviceToHost) != hipSuccess) return RX_ERR_DEVICE;
  b.fix_overflow = b.last_mode == 1 && ctr[CTR_FIX_OVF] != 0;
  b.counters[0] = std::min(ctr[CTR_ITEMS] + ctr[CTR_ITEMS_S], b.item_cap);
  b.counters[1] = (b.last_mode == 1) ? ctr[CTR_FIX] : 0;
  b.counters[2] = b.fix_overflow ? 1 : 0;
  if (b.fix_overflow) {
    // the FP64 re-check queue overflowed: redo the batch in FP64
    std::string err;
    if (batch_run(b, RX_FP64, err) != RX_OK) return RX_ERR_DEVICE;
    if (hipStreamSynchronize(b.stream) != hipSuccess) return RX_ERR_DEVICE;
  }
  if (b.profiled) {
    // events: setup | integrate | fixup | finalize | xill | conv | primary; the
    // timing slots keep conv fourth and add xill and primary after it
    static const int slot[RX_N_TIMED] = {0, 1, 2, 3, 5, 4, 6};
    for (int k = 0; k < RX_N_TIMED; k++) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, b.ev[k], b.ev[k + 1]);
      b.kernel_ms[slot[k]] = ms;
    }
  }
  return RX_OK;
}

int public_status(int dev_status) {
  if (dev_status == ST_WORKLIST) return RX_ERR_DEVICE;
  if (dev_status >= 10 && dev_status < 50) return RX_ERR_TABLE;
  return dev_status;
}

}  // namespace rx
