// capi.cpp -- extern "C" entry points of librelxill_gpu.so.
//
// Nothing throws across this boundary: every error becomes a status (the
// reference lets ModelEvalFailed escape its extern "C" wrapper,
// src/LocalModel.cpp:143-160).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>

#include "engine.h"
#include "rx_math.h"

using namespace rx;

struct rx_batch {
  Batch b;
};

namespace {

// one reusable single-request workspace for the XSPEC entry points
Batch &plain_batch() {
  static Batch b;
  return b;
}

int eval_one(int model, const double *parameter, const double *energy, double *flux, int n_flux,
             int precision) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  std::string err;
  if (!energy || !parameter || !flux || n_flux <= 0) {
    set_last(RX_ERR_ARG, "null pointer or empty grid");
    return RX_ERR_ARG;
  }
  int st = ensure_device(err);
  if (st != RX_OK) {
    for (int i = 0; i < n_flux; i++) flux[i] = 0.0;
    set_last(st, err);
    return st;
  }
  rx_request r;
  r.model = model;
  r.n_flux = n_flux;
  r.energy = energy;
  r.parameter = parameter;
  r.flux_in = flux;
  r.refl_set = -1;
  Batch &b = plain_batch();
  st = batch_build(b, &r, 1, err);
  if (st == RX_OK) st = batch_run(b, precision, err);
  if (st == RX_OK) st = batch_sync(b);
  if (st == RX_OK) {
    if (hipMemcpy(flux, b.d_out.p, sizeof(double) * n_flux, hipMemcpyDeviceToHost) != hipSuccess) {
      st = RX_ERR_DEVICE;
      err = "device to host copy failed";
    }
  }
  if (st == RX_OK) {
    int dst = 0;
    if (hipMemcpy(&dst, b.d_status.p, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) dst = RX_ERR_DEVICE;
    if (dst != 0) {
      st = public_status(dst);
      err = "model evaluation failed (status " + std::to_string(dst) + ")";
    }
  }
  if (st != RX_OK) {
    for (int i = 0; i < n_flux; i++) flux[i] = 0.0;
  }
  set_last(st, err);
  return st;
}

}  // namespace

extern "C" {

int rx_eval_model(int model, const double *parameter, const double *energy, double *flux, int n_flux,
                  int precision) {
  try {
    return eval_one(model, parameter, energy, flux, n_flux, precision);
  } catch (const std::bad_alloc &) {
    set_last(RX_ERR_NOMEM, "out of host memory");
    return RX_ERR_NOMEM;
  } catch (...) {
    set_last(RX_ERR_DEVICE, "unexpected exception");
    return RX_ERR_DEVICE;
  }
}

void lmodrelline(const double *energy, int Nflux, const double *parameter, int, double *flux, double *,
                 const char *) {
  rx_eval_model(RX_MODEL_RELLINE, parameter, energy, flux, Nflux, RX_FP64);
}

void lmodrellinelp(const double *energy, int Nflux, const double *parameter, int, double *flux, double *,
                   const char *) {
  rx_eval_model(RX_MODEL_RELLINELP, parameter, energy, flux, Nflux, RX_FP64);
}

void lmodrelconv(const double *energy, int Nflux, const double *parameter, int, double *flux, double *,
                 const char *) {
  rx_eval_model(RX_MODEL_RELCONV, parameter, energy, flux, Nflux, RX_FP64);
}

void lmodrelconvlp(const double *energy, int Nflux, const double *parameter, int, double *flux, double *,
                   const char *) {
  rx_eval_model(RX_MODEL_RELCONVLP, parameter, energy, flux, Nflux, RX_FP64);
}

void lmodrelxill(const double *energy, int Nflux, const double *parameter, int, double *flux, double *,
                 const char *) {
  rx_eval_model(RX_MODEL_RELXILL, parameter, energy, flux, Nflux, RX_FP64);
}

void lmodrelxilllp(const double *energy, int Nflux, const double *parameter, int, double *flux, double *,
                   const char *) {
  rx_eval_model(RX_MODEL_RELXILLLP, parameter, energy, flux, Nflux, RX_FP64);
}

void lmodrelxilldensnthcomp(const double *energy, int Nflux, const double *parameter, int, double *flux, double *,
                            const char *) {
  rx_eval_model(RX_MODEL_RELXILLCP, parameter, energy, flux, Nflux, RX_FP64);
}

void lmodrelxilllpdensnthcomp(const double *energy, int Nflux, const double *parameter, int, double *flux,
                              double *, const char *) {
  rx_eval_model(RX_MODEL_RELXILLLPCP, parameter, energy, flux, Nflux, RX_FP64);
}

void lmodrelxillns(const double *energy, int Nflux, const double *parameter, int, double *flux, double *,
                   const char *) {
  rx_eval_model(RX_MODEL_RELXILLNS, parameter, energy, flux, Nflux, RX_FP64);
}

void lmodxillver(const double *energy, int Nflux, const double *parameter, int, double *flux, double *,
                 const char *) {
  rx_eval_model(RX_MODEL_XILLVER, parameter, energy, flux, Nflux, RX_FP64);
}

void lmodxillverdensnthcomp(const double *energy, int Nflux, const double *parameter, int, double *flux, double *,
                            const char *) {
  rx_eval_model(RX_MODEL_XILLVERCP, parameter, energy, flux, Nflux, RX_FP64);
}

void lmodxillverns(const double *energy, int Nflux, const double *parameter, int, double *flux, double *,
                   const char *) {
  rx_eval_model(RX_MODEL_XILLVERNS, parameter, energy, flux, Nflux, RX_FP64);
}

int rx_last_status(void) { return last_status(); }
const char *rx_last_error(void) { return last_error(); }

int rx_init(int device) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (g.device_ready && g.device != device) {
    set_last(RX_ERR_ARG, "library already initialised on another device");
    return RX_ERR_ARG;
  }
  g.device = device;
  std::string err;
  int st = ensure_device(err);
  set_last(st, err);
  return st;
}

void rx_shutdown(void) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (g.device_ready) (void)hipDeviceSynchronize();
}

int rx_tables_synthetic(uint64_t seed, int n_a, int n_mu0, int n_r, int n_g) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (n_a < 2 || n_mu0 < 2 || n_r < 3 || n_g < 2 || n_r > MAX_NR_TAB || n_g > MAX_NG) {
    set_last(RX_ERR_ARG, "table dims out of range");
    return RX_ERR_ARG;
  }
  try {
    g.rel = make_synthetic_reltable(seed, n_a, n_mu0, n_r, n_g);
  } catch (...) {
    set_last(RX_ERR_NOMEM, "table allocation failed");
    return RX_ERR_NOMEM;
  }
  g.rel_loaded = true;
  g.rel_err.clear();
  g.rel_dirty = true;
  g.config_gen++;  // batches built before hold buffers sized for the old table
  std::string e = validate_reltable(g.rel);
  set_last(e.empty() ? RX_OK : RX_ERR_TABLE, e);
  return e.empty() ? RX_OK : RX_ERR_TABLE;
}

int rx_lptable_synthetic(uint64_t seed, int n_a, int n_h, int n_r) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (n_a < 2 || n_h < 2 || n_r < 2 || n_r > MAX_LP_R) {
    set_last(RX_ERR_ARG, "lamp-post table dims out of range");
    return RX_ERR_ARG;
  }
  g.lp = make_synthetic_lptable(seed, n_a, n_h, n_r);
  g.lp_loaded = true;
  g.lp_err.clear();
  g.lp_dirty = true;
  g.config_gen++;
  std::string e = validate_lptable(g.lp);
  set_last(e.empty() ? RX_OK : RX_ERR_TABLE, e);
  return e.empty() ? RX_OK : RX_ERR_TABLE;
}

int rx_tables_load(const char *path) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!path) return RX_ERR_ARG;
  RelTableHost t;
  std::string e = load_reltable(t, path);
  if (!e.empty()) {
    set_last(RX_ERR_TABLE, e);
    return RX_ERR_TABLE;
  }
  g.rel = std::move(t);
  g.rel_loaded = true;
  g.rel_err.clear();
  g.rel_dirty = true;
  g.config_gen++;  // batches built before hold buffers sized for the old table
  return RX_OK;
}

int rx_lptable_load(const char *path) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!path) return RX_ERR_ARG;
  LpTableHost t;
  std::string e = load_lptable(t, path);
  if (!e.empty()) {
    set_last(RX_ERR_TABLE, e);
    return RX_ERR_TABLE;
  }
  g.lp = std::move(t);
  g.lp_loaded = true;
  g.lp_err.clear();
  g.lp_dirty = true;
  g.config_gen++;
  return RX_OK;
}

int rx_tables_load_fits(const char *path) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!path) return RX_ERR_ARG;
  RelTableHost t;
  std::string e = load_reltable_fits(t, path);
  if (!e.empty()) {
    set_last(RX_ERR_TABLE, e);
    return RX_ERR_TABLE;
  }
  g.rel = std::move(t);
  g.rel_loaded = true;
  g.rel_err.clear();
  g.rel_dirty = true;
  g.config_gen++;
  return RX_OK;
}

int rx_lptable_load_fits(const char *path) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!path) return RX_ERR_ARG;
  LpTableHost t;
  std::string e = load_lptable_fits(t, path);
  if (!e.empty()) {
    set_last(RX_ERR_TABLE, e);
    return RX_ERR_TABLE;
  }
  g.lp = std::move(t);
  g.lp_loaded = true;
  g.lp_err.clear();
  g.lp_dirty = true;
  g.config_gen++;
  return RX_OK;
}

static int xill_synthetic(int prim, uint64_t seed, int n_gam, int n_afe, int n_lxi, int n_ect, int n_incl,
                          int n_ener) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (n_gam < 2 || n_afe < 2 || n_lxi < 2 || n_ect < 2 || n_incl < 2 || n_incl > MAX_INCL || n_ener < 2) {
    set_last(RX_ERR_ARG, "xillver table dims out of range");
    return RX_ERR_ARG;
  }
  XillSlot &x = g.xill[prim];
  try {
    // the fourth axis: Ecut 5-1000 keV, or kTe 1-400 keV for nthcomp; the
    // blackbody table's axes are kTbb, A_Fe, logXi, Dens, Incl
    x.t = prim == PRIM_NTHCOMP ? make_synthetic_xilltable(seed, n_gam, n_afe, n_lxi, n_ect, n_incl, n_ener, 1.0, 400.0)
          : prim == PRIM_BBODY ? make_synthetic_xilltable_ns(seed, n_gam, n_afe, n_lxi, n_ect, n_incl, n_ener)
                               : make_synthetic_xilltable(seed, n_gam, n_afe, n_lxi, n_ect, n_incl, n_ener);
  } catch (...) {
    set_last(RX_ERR_NOMEM, "table allocation failed");
    return RX_ERR_NOMEM;
  }
  x.loaded = true;
  x.err.clear();
  x.dirty = true;
  g.config_gen++;
  std::string e = validate_xilltable(x.t);
  set_last(e.empty() ? RX_OK : RX_ERR_TABLE, e);
  return e.empty() ? RX_OK : RX_ERR_TABLE;
}

static int xill_load_fits(int prim, const char *path) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!path) return RX_ERR_ARG;
  XillTableHost t;
  std::string e;
  try {
    e = load_xilltable_fits(t, path);
  } catch (...) {
    e = "table allocation failed";
  }
  if (!e.empty()) {
    set_last(RX_ERR_TABLE, e);
    return RX_ERR_TABLE;
  }
  if (t.n_incl() > MAX_INCL) {
    set_last(RX_ERR_TABLE, "xillver table has more inclinations than the kernels support");
    return RX_ERR_TABLE;
  }
  XillSlot &x = g.xill[prim];
  x.t = std::move(t);
  x.loaded = true;
  x.err.clear();
  x.dirty = true;
  g.config_gen++;
  return RX_OK;
}

static int xill_get_view(int prim, rx_xilltable_view *v) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!v) return RX_ERR_ARG;
  ensure_xilltable_host(prim);
  const XillSlot &x = g.xill[prim];
  if (!x.loaded) {
    set_last(RX_ERR_TABLE, "xillver table: " + x.err);
    return RX_ERR_TABLE;
  }
  const XillTableHost &t = x.t;
  memset(v, 0, sizeof(*v));
  v->n_par = t.n_par;
  for (int k = 0; k < t.n_par && k < 6; k++) {
    v->pid[k] = t.pid[k];
    v->n_val[k] = t.n_val[k];
    v->val[k] = t.val[k].data();
  }
  v->n_ener = t.n_ener;
  v->elo = t.elo.data();
  v->ehi = t.ehi.data();
  v->spec = t.spec.data();
  return RX_OK;
}

int rx_xilltable_synthetic(uint64_t seed, int n_gam, int n_afe, int n_lxi, int n_ect, int n_incl, int n_ener) {
  return xill_synthetic(PRIM_CUTOFFPL, seed, n_gam, n_afe, n_lxi, n_ect, n_incl, n_ener);
}
int rx_xilltable_load_fits(const char *path) { return xill_load_fits(PRIM_CUTOFFPL, path); }
int rx_xilltable_get_view(rx_xilltable_view *v) { return xill_get_view(PRIM_CUTOFFPL, v); }

int rx_xillcptable_synthetic(uint64_t seed, int n_gam, int n_afe, int n_lxi, int n_kte, int n_incl, int n_ener) {
  return xill_synthetic(PRIM_NTHCOMP, seed, n_gam, n_afe, n_lxi, n_kte, n_incl, n_ener);
}
int rx_xillcptable_synthetic_dens(uint64_t seed, int n_dens, int n_gam, int n_afe, int n_lxi, int n_kte, int n_incl,
                                  int n_ener) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (n_dens < 2 || n_gam < 2 || n_afe < 2 || n_lxi < 2 || n_kte < 2 || n_incl < 2 || n_incl > MAX_INCL ||
      n_ener < 2) {
    set_last(RX_ERR_ARG, "table dimensions out of range");
    return RX_ERR_ARG;
  }
  XillSlot &x = g.xill[PRIM_NTHCOMP];
  try {
    x.t = make_synthetic_xilltable_dens(seed, n_dens, n_gam, n_afe, n_lxi, n_kte, n_incl, n_ener, 1.0, 400.0);
  } catch (...) {
    set_last(RX_ERR_NOMEM, "table allocation failed");
    return RX_ERR_NOMEM;
  }
  x.loaded = true;
  x.err.clear();
  x.dirty = true;
  g.config_gen++;
  std::string e = validate_xilltable(x.t);
  set_last(e.empty() ? RX_OK : RX_ERR_TABLE, e);
  return e.empty() ? RX_OK : RX_ERR_TABLE;
}
int rx_xillcptable_load_fits(const char *path) { return xill_load_fits(PRIM_NTHCOMP, path); }
int rx_xillcptable_get_view(rx_xilltable_view *v) { return xill_get_view(PRIM_NTHCOMP, v); }

int rx_xillnstable_synthetic(uint64_t seed, int n_ktbb, int n_afe, int n_lxi, int n_dens, int n_incl, int n_ener) {
  return xill_synthetic(PRIM_BBODY, seed, n_ktbb, n_afe, n_lxi, n_dens, n_incl, n_ener);
}
int rx_xillnstable_load_fits(const char *path) { return xill_load_fits(PRIM_BBODY, path); }
int rx_xillnstable_get_view(rx_xilltable_view *v) { return xill_get_view(PRIM_BBODY, v); }

int rx_rradtable_synthetic(uint64_t seed, int n_spin, int n_rad, int n_g) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (n_spin < 1 || n_rad < 2 || n_g < 1 || n_rad > 256 || n_g > 256) {
    set_last(RX_ERR_ARG, "returning-radiation table dims out of range");
    return RX_ERR_ARG;
  }
  g.rrad = make_synthetic_rradtable(seed, n_spin, n_rad, n_g);
  g.rrad_loaded = true;
  g.rrad_err.clear();
  g.rrad_dirty = true;
  g.config_gen++;
  std::string e = validate_rradtable(g.rrad);
  set_last(e.empty() ? RX_OK : RX_ERR_TABLE, e);
  return e.empty() ? RX_OK : RX_ERR_TABLE;
}

int rx_rradtable_load_fits(const char *path) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!path) return RX_ERR_ARG;
  RradTableHost t;
  std::string e = load_rradtable_fits(t, path);
  if (!e.empty()) {
    set_last(RX_ERR_TABLE, e);
    return RX_ERR_TABLE;
  }
  g.rrad = std::move(t);
  g.rrad_loaded = true;
  g.rrad_err.clear();
  g.rrad_dirty = true;
  g.config_gen++;
  return RX_OK;
}

int rx_rradtable_get_view(rx_rradtable_view *v) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!v) return RX_ERR_ARG;
  ensure_rradtable_host();
  if (!g.rrad_loaded) {
    set_last(RX_ERR_TABLE, "returning-radiation table: " + g.rrad_err);
    return RX_ERR_TABLE;
  }
  const RradTableHost &t = g.rrad;
  *v = {t.n_spin, t.n_rad, t.n_g, t.spin.data(), t.rlo.data(), t.rhi.data(), t.tf_r.data(), t.gmin.data(),
        t.gmax.data(), t.frac_g.data()};
  return RX_OK;
}

int rx_tables_save(const char *path) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  ensure_tables_host();
  if (!g.rel_loaded) {
    set_last(RX_ERR_TABLE, g.rel_err);
    return RX_ERR_TABLE;
  }
  std::string e = save_reltable(g.rel, path ? path : "");
  set_last(e.empty() ? RX_OK : RX_ERR_ARG, e);
  return e.empty() ? RX_OK : RX_ERR_ARG;
}

int rx_lptable_save(const char *path) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  ensure_tables_host();
  if (!g.lp_loaded) {
    set_last(RX_ERR_TABLE, g.lp_err);
    return RX_ERR_TABLE;
  }
  std::string e = save_lptable(g.lp, path ? path : "");
  set_last(e.empty() ? RX_OK : RX_ERR_ARG, e);
  return e.empty() ? RX_OK : RX_ERR_ARG;
}

int rx_tables_view(rx_reltable_view *rel, rx_lptable_view *lp) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  ensure_tables_host();
  if ((rel && !g.rel_loaded) || (lp && !g.lp_loaded)) {
    set_last(RX_ERR_TABLE, !g.rel_loaded ? g.rel_err : g.lp_err);
    return RX_ERR_TABLE;
  }
  if (rel) {
    const RelTableHost &t = g.rel;
    *rel = {t.n_a, t.n_mu0, t.n_r, t.n_g, t.a.data(), t.mu0.data(), t.r.data(), t.gmin.data(),
            t.gmax.data(), t.trff1.data(), t.trff2.data(), t.cosne1.data(), t.cosne2.data()};
  }
  if (lp) {
    const LpTableHost &t = g.lp;
    *lp = {t.n_a, t.n_h, t.n_r, t.a.data(), t.h.data(), t.rad.data(), t.intens.data(), t.del.data(),
           t.del_inc.data()};
  }
  return RX_OK;
}

int rx_set_conv_grid(int n, double emin, double emax) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (n < 256 || n > MAX_FFT || (n & (n - 1)) != 0 || !(emin > 0) || !(emax > emin)) {
    set_last(RX_ERR_ARG, "convolution grid must have a power-of-two size in [256, 8192]");
    return RX_ERR_ARG;
  }
  if (!(emin < 1.0 && emax > 1.0)) {
    set_last(RX_ERR_ARG, "convolution grid must contain 1 keV");
    return RX_ERR_ARG;
  }
  if (!g.refl_sets.empty() && n != g.conv_n) {
    set_last(RX_ERR_ARG, "cannot change the convolution grid size with resident reflection sets");
    return RX_ERR_ARG;
  }
  g.conv_n = n;
  g.conv_emin = emin;
  g.conv_emax = emax;
  g.conv_dirty = true;
  g.config_gen++;
  ensure_tables_host();
  return RX_OK;
}

int rx_set_nr_fine(int nr_fine) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (nr_fine < 3 || nr_fine > MAX_NR_FINE) return RX_ERR_ARG;
  if (nr_fine != g.nr_fine) g.config_gen++;
  g.nr_fine = nr_fine;
  return RX_OK;
}

/* ---- batches ---- */

rx_batch *rx_batch_create(const rx_request *req, int n_req) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  std::string err;
  try {
    int st = ensure_device(err);
    if (st != RX_OK) {
      set_last(st, err);
      return nullptr;
    }
    std::unique_ptr<rx_batch> b(new rx_batch);
    st = batch_build(b->b, req, n_req, err);
    if (st != RX_OK) {
      set_last(st, err);
      return nullptr;
    }
    set_last(RX_OK, "");
    return b.release();
  } catch (...) {
    set_last(RX_ERR_NOMEM, "batch allocation failed");
    return nullptr;
  }
}

int rx_batch_run(rx_batch *b, int precision) {
  if (!b) return RX_ERR_ARG;
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  std::string err;
  int st = ensure_device(err);
  if (st == RX_OK) st = batch_run(b->b, precision, err);
  set_last(st, err);
  return st;
}

int rx_batch_sync(rx_batch *b) {
  if (!b) return RX_ERR_ARG;
  return batch_sync(b->b);
}

int rx_batch_status(rx_batch *b, int i) {
  if (!b || i < 0 || i >= b->b.n_req) return RX_ERR_ARG;
  int st = 0;
  if (hipMemcpy(&st, b->b.d_status.as<int>() + i, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return RX_ERR_DEVICE;
  return public_status(st);
}

int rx_batch_output(rx_batch *b, int i, double *flux) {
  if (!b || !flux || i < 0 || i >= b->b.n_req) return RX_ERR_ARG;
  const ReqParams &p = b->b.h_req[i];
  if (hipMemcpy(flux, b->b.d_out.as<double>() + p.flux_off, sizeof(double) * p.n_in,
                hipMemcpyDeviceToHost) != hipSuccess)
    return RX_ERR_DEVICE;
  return rx_batch_status(b, i);
}

int rx_batch_dist(rx_batch *b, int i, double *dist, int n_incl) {
  if (!b || !dist || i < 0 || i >= b->b.n_req || n_incl < 0) return RX_ERR_ARG;
  const ReqParams &p = b->b.h_req[i];
  if (n_incl > p.n_incl * p.n_zones) return RX_ERR_ARG;
  if (n_incl == 0) return RX_OK;
  if (hipMemcpy(dist, b->b.d_dist.as<double>() + p.dist_off, sizeof(double) * n_incl,
                hipMemcpyDeviceToHost) != hipSuccess)
    return RX_ERR_DEVICE;
  return RX_OK;
}

int rx_batch_zones(rx_batch *b, int i, double *out, int n_zones) {
  if (!b || !out || i < 0 || i >= b->b.n_req || n_zones < 0) return RX_ERR_ARG;
  const ReqParams &p = b->b.h_req[i];
  if (p.model != MOD_RELXILL && p.model != MOD_RELXILLLP) return RX_ERR_ARG;
  if (n_zones > p.n_zones) return RX_ERR_ARG;
  if (n_zones == 0) return RX_OK;
  std::vector<ZoneRec> z(n_zones);
  if (hipMemcpy(z.data(), b->b.d_zone.as<ZoneRec>() + p.zone_off, sizeof(ZoneRec) * n_zones,
                hipMemcpyDeviceToHost) != hipSuccess)
    return RX_ERR_DEVICE;
  for (int k = 0; k < n_zones; k++) {
    double *o = out + 8 * k;
    o[0] = z[k].rmean;
    o[1] = z[k].del_emit;
    o[2] = z[k].shift;
    o[3] = z[k].ncf;
    o[4] = z[k].ect_ind;
    o[5] = z[k].ect_fac;
    o[6] = z[k].lxi_ind;
    o[7] = z[k].lxi_fac;
  }
  return RX_OK;
}

int rx_batch_reflfrac(rx_batch *b, int i, double *out5) {
  if (!b || !out5 || i < 0 || i >= b->b.n_req) return RX_ERR_ARG;
  LpRefl r;
  if (hipMemcpy(&r, b->b.d_refl.as<LpRefl>() + i, sizeof(LpRefl), hipMemcpyDeviceToHost) != hipSuccess)
    return RX_ERR_DEVICE;
  out5[0] = r.refl_frac;
  out5[1] = r.f_bh;
  out5[2] = r.f_ad;
  out5[3] = r.f_inf;
  out5[4] = r.f_inf_rest;
  return RX_OK;
}

int rx_batch_set_shape(rx_batch *b, int seg_bins, int waves_per_cu) {
  if (!b || seg_bins < 0 || seg_bins > 4096 || waves_per_cu < 0 || waves_per_cu > 32) return RX_ERR_ARG;
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  b->b.seg_forced = seg_bins;
  b->b.waves_forced = waves_per_cu;
  batch_plan(b->b, seg_bins, waves_per_cu);
  return batch_upload(b->b);
}

void rx_batch_destroy(rx_batch *b) {
  if (!b) return;
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (b->b.stream) (void)hipStreamSynchronize(b->b.stream);
  delete b;
}

int rx_copy_test(int64_t bytes, int reps, double *ms) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  std::string err;
  int st = ensure_device(err);
  if (st != RX_OK || bytes <= 0 || reps <= 0) return st != RX_OK ? st : RX_ERR_ARG;
  DevBuf a, b;
  if (a.ensure((size_t)bytes) || b.ensure((size_t)bytes)) return RX_ERR_NOMEM;
  (void)hipMemset(a.p, 1, (size_t)bytes);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  int e = launch_copy(a.p, b.p, (size_t)bytes, nullptr);
  (void)hipEventRecord(e0, nullptr);
  for (int r = 0; r < reps; r++) e |= launch_copy(a.p, b.p, (size_t)bytes, nullptr);
  (void)hipEventRecord(e1, nullptr);
  (void)hipEventSynchronize(e1);
  float t = 0;
  (void)hipEventElapsedTime(&t, e0, e1);
  if (ms) *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return e ? RX_ERR_DEVICE : RX_OK;
}

int rx_nthcomp(const double *energy, int n_flux, int n_spec, const double *kte, const double *gam,
               const double *ener_shift, double *flux) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!energy || !kte || !gam || !ener_shift || !flux || n_flux <= 0 || n_spec <= 0) {
    set_last(RX_ERR_ARG, "null pointer, empty grid or no spectrum");
    return RX_ERR_ARG;
  }
  std::string err;
  int st = ensure_device(err);
  if (st == RX_OK) {
    try {
      st = nthcomp_spectra(energy, n_flux, n_spec, kte, gam, ener_shift, flux, err);
    } catch (...) {
      st = RX_ERR_DEVICE;
      err = "unexpected exception";
    }
  }
  set_last(st, err);
  return st;
}

int rx_device_synchronize(void) {
  Global &g = G();
  if (!g.device_ready) return RX_OK;
  return hipDeviceSynchronize() == hipSuccess ? RX_OK : RX_ERR_DEVICE;
}

int rx_set_fp32_guard(double rel) {
  if (!(rel >= 0.0 && rel < 1.0)) return RX_ERR_ARG;
  G().guard = (float)rel;
  return RX_OK;
}

int rx_profile_enable(int on) {
  G().profile = (on != 0);
  return RX_OK;
}

int rx_batch_kernel_ms(rx_batch *b, double *ms) {
  if (!b || !ms) return RX_ERR_ARG;
  for (int k = 0; k < RX_N_TIMED; k++) ms[k] = b->b.kernel_ms[k];
  return b->b.profiled ? RX_OK : RX_ERR_ARG;
}

int rx_set_graphs(int on) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  g.use_graphs = on != 0;
  return RX_OK;
}

int64_t rx_batch_graph_replays(rx_batch *b) {
  if (!b) return -RX_ERR_ARG;
  return b->b.graph_replays;
}

int rx_batch_counters(rx_batch *b, int64_t *out) {
  if (!b || !out) return RX_ERR_ARG;
  for (int k = 0; k < 3; k++) out[k] = b->b.counters[k];
  return RX_OK;
}

// Times every (segment bins, waves per CU) candidate on this batch and keeps
// the fastest for batches of the same shape class (log2 of the profile bins
// and whether convolutions are present).
int rx_batch_autotune(rx_batch *b, int precision) {
  if (!b) return RX_ERR_ARG;
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  std::string err;
  const int segs[] = {32, 64, 128, 256};
  const int wpcs[] = {4, 8, 12, 16};
  struct Cand {
    int seg, wpc;
    double dt;
  };
  std::vector<Cand> cands;
  // the fastest of `trials` timings of `reps` runs each of one shape
  auto time_shape = [&](int sg, int w, int trials, int reps, double &dt) {
    batch_plan(b->b, sg, w);
    if (batch_upload(b->b) != RX_OK) return false;
    if (batch_run(b->b, precision, err) != RX_OK || batch_sync(b->b) != RX_OK) return false;
    for (int trial = 0; trial < trials; trial++) {
      auto t0 = std::chrono::steady_clock::now();
      for (int k = 0; k < reps; k++) {
        if (batch_run(b->b, precision, err) != RX_OK) return false;
      }
      if (batch_sync(b->b) != RX_OK) return false;
      dt = std::min(dt, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / reps);
    }
    return true;
  };
  for (int w : wpcs) {
    for (int sg : segs) {
      double dt = 1e30;
      if (!time_shape(sg, w, 3, 2, dt)) return -RX_ERR_DEVICE;
      cands.push_back({sg, w, dt});
    }
  }
  // the three fastest timed again, in turns, with more runs: shapes within a
  // few percent of each other are told apart despite the host timer's noise
  std::sort(cands.begin(), cands.end(), [](const Cand &x, const Cand &y) { return x.dt < y.dt; });
  cands.resize(std::min<size_t>(3, cands.size()));
  for (auto &c : cands) c.dt = 1e30;
  for (int round = 0; round < 3; round++)
    for (auto &c : cands)
      if (!time_shape(c.seg, c.wpc, 2, 3, c.dt)) return -RX_ERR_DEVICE;
  std::sort(cands.begin(), cands.end(), [](const Cand &x, const Cand &y) { return x.dt < y.dt; });
  const int best_s = cands[0].seg, best_w = cands[0].wpc;
  bool found = false;
  for (auto &t : g.tune)
    if (t.shape == b->b.shape_class) {
      t.seg_bins = best_s;
      t.waves_per_cu = best_w;
      found = true;
    }
  if (!found) g.tune.push_back({b->b.shape_class, best_s, best_w});
  b->b.seg_forced = best_s;
  b->b.waves_forced = best_w;
  batch_plan(b->b, best_s, best_w);
  if (batch_upload(b->b) != RX_OK) return -RX_ERR_DEVICE;
  return best_s * 256 + best_w;
}

/* ---- reflection sets ---- */

int rx_reflset_create(const double *refl_energy, int n_refl, const double *refl, int n_incl) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (!refl_energy || !refl || n_refl <= 0 || n_incl <= 0 || n_incl > MAX_INCL) {
    set_last(RX_ERR_ARG, "invalid reflection set");
    return -RX_ERR_ARG;
  }
  std::string err;
  int st = ensure_device(err);
  if (st != RX_OK) {
    set_last(st, err);
    return -st;
  }
  const int n = g.conv_n;
  // buffers: energy, spectra, rebinned spectra, band sums, FFTs
  DevBuf *bufs = new DevBuf[5];
  int e = bufs[0].ensure(sizeof(double) * (n_refl + 1)) | bufs[1].ensure(sizeof(double) * (size_t)n_refl * n_incl) |
          bufs[2].ensure(sizeof(double) * (size_t)n * n_incl) | bufs[3].ensure(sizeof(double) * n_incl) |
          bufs[4].ensure(sizeof(double) * (size_t)(n / 2 + 1) * 2 * n_incl);
  if (e) {
    delete[] bufs;
    set_last(RX_ERR_NOMEM, "device allocation failed");
    return -RX_ERR_NOMEM;
  }
  e |= (int)hipMemcpy(bufs[0].p, refl_energy, sizeof(double) * (n_refl + 1), hipMemcpyHostToDevice);
  e |= (int)hipMemcpy(bufs[1].p, refl, sizeof(double) * (size_t)n_refl * n_incl, hipMemcpyHostToDevice);
  BatchDev bd;
  memset(&bd, 0, sizeof(bd));
  fill_dev_views(bd);
  e |= launch_refl_prepare(bd.cg, bufs[0].as<double>(), bufs[1].as<double>(), n_refl, n_incl,
                           bufs[2].as<double>(), bufs[3].as<double>(), bufs[4].as<double>(), nullptr);
  e |= (int)hipDeviceSynchronize();
  if (e) {
    delete[] bufs;
    set_last(RX_ERR_DEVICE, "reflection set preparation failed");
    return -RX_ERR_DEVICE;
  }
  DevReflSet rs;
  rs.n_incl = n_incl;
  rs.spec = bufs[2].as<double>();
  rs.bandsum = bufs[3].as<double>();
  rs.fft = bufs[4].as<double>();
  g.refl_sets.push_back(rs);
  g.refl_bufs.push_back(bufs);
  g.refl_dirty = true;
  st = ensure_device(err);
  if (st != RX_OK) return -st;
  return (int)g.refl_sets.size() - 1;
}

int rx_reflsets_clear(void) {
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  if (g.device_ready) (void)hipDeviceSynchronize();  // no running batch may still read them
  for (DevBuf *bufs : g.refl_bufs) delete[] bufs;
  g.refl_bufs.clear();
  g.refl_sets.clear();
  g.refl_dirty = true;
  g.config_gen++;  // set ids of existing batches are void
  set_last(RX_OK, "");
  return RX_OK;
}

/* ---- parameter processing (host mirror of get_rel_params), for checking ---- */
int rx_model_params(int model, const double *parameter, double *out12) {
  if (!parameter || !out12) return RX_ERR_ARG;
  ReqParams p;
  std::string err;
  int st = parse_request(model, parameter, phys_norm_env(), p, err);
  const double v[12] = {p.a, p.incl, p.emis1, p.emis2, p.rbr, p.rin, p.rout, p.lineE, p.z, p.height,
                        p.gamma, p.beta};
  memcpy(out12, v, sizeof v);
  set_last(st, err);
  return st;
}

}  // extern "C"
