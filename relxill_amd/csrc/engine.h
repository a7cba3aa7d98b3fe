// engine.h -- host runtime of librelxill_gpu (internal).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/relxill_gpu.h"
#include "rx_engine.h"
#include "rx_math.h"
#include "rx_tables.h"

namespace rx {

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  ~DevBuf();
  int ensure(size_t n);  // grows (never shrinks); returns hipError_t
  template <class T>
  T *as() const { return reinterpret_cast<T *>(p); }
};

// Host mirror of get_rel_params + check_parameter_bounds
// (src/ModelDefinition.cpp:153-334) for one request.  Returns RX_OK or a status.
// prim: the primary continuum of the relxill models (PRIM_*)
int parse_request(int model, const double *par, bool phys_norm, ReqParams &p, std::string &err,
                  int prim = PRIM_CUTOFFPL);
// relxill requests: the xillver table corner and factors of the request's
// parameters (get_xilltab_indices_for_paramvals and interp_xill_table,
// src/xilltable.c:297-324, :1090-1125).
// ind / fac: the four inner axes.  The xillver models interpolate the
// inclination too (clamped into the table like the others): clamp_incl, and
// its bracket and factor in *inc_ind / *inc_fac.  A six-parameter table's
// first axis (interpolated outside the others) goes to *out_ind / *out_fac
// (0 and 0 for five parameters).
void xill_corner(const XillTableHost &t, double gam, double afe, double lxi, double ect, double dens, double ktbb,
                 double incl_deg, int *ind, double *fac, bool clamp_incl = false, int *inc_ind = nullptr,
                 double *inc_fac = nullptr, int *out_ind = nullptr, double *out_fac = nullptr);
int n_model_params(int model);
// Edges of the radial zones (RadialGrid::calculate_radial_grid, src/IonGradient.cpp:457-489).
void rzone_grid(double rmin, double rmax, int nzones, double h, double *rgrid);

// Launch configuration of one run: every kernel argument and grid shape.  A
// run whose configuration equals the previous run's is captured into a
// hipGraph, which is replayed for as long as the configuration stays equal
// (the kernels read the batch's data through the same pointers).  Built with
// its padding zeroed, so configurations compare with memcmp.
struct RunKey {
  BatchDev bd;
  int n_blocks, n_xill1, n_xill, n_conv, n_prim, n_rrc, n_nth;
  const void *d_xill, *d_conv, *d_prim, *d_rrc, *d_nth;
  int prof, cos_rows;
};

struct Batch {
  int n_req = 0;
  int max_zones = 1;          // most zones of a request
  int seg_bins = 64;        // bins per work item
  int seg_forced = 0;       // 0: automatic / autotuned
  int waves_per_cu = 8;     // persistent integration wavefronts per CU
  int waves_forced = 0;
  std::vector<ReqParams> h_req;
  std::vector<int> h_status;  // host-side parse status
  std::vector<double> h_ener, h_flux_in;
  std::vector<int> h_prim;
  std::vector<XillWork> h_conv;  // (request, zone) units of k_conv
  std::vector<XillWork> h_xill;  // k_xill units: one-zone requests first, then the others (all zones in one unit)
  int n_xill1 = 0;               // number of one-zone units at the front of h_xill
  std::vector<XillWork> h_rrc;   // (request, zone) units of the returning-radiation corrections
  std::vector<XillWork> h_nth;   // nthcomp solves: (request, 0 source / 1 + zone)
  int n_nth_sol = 0;             // requests with the nthcomp primary (one source solution each)
  int64_t tot_ener = 0, tot_flux = 0, tot_pgrid = 0, tot_prof = 0, tot_xspec = 0, tot_dist = 0, tot_zone = 0;
  int64_t tot_node = 0;  // node rows of the requests with an angle distribution
  int64_t tot_rrc = 0;
  DevBuf d_req, d_ener, d_flux_in, d_out, d_pgrid, d_node, d_rad, d_prof, d_dist, d_refl, d_status, d_conv;
  DevBuf d_xill, d_xspec, d_zone, d_prim, d_rrc, d_rrcorr;
  DevBuf d_nthw, d_nth, d_nthx, d_nths, d_nthn;
  DevBuf d_items, d_ctr, d_fix, d_reach, d_rnorm, d_rnode;
  int item_cap = 0;
  bool any_cosne = false, any_limb = false;
  int fix_cap = 0;
  int last_mode = 1;
  bool fix_overflow = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev[RX_N_TIMED + 1] = {};
  bool profiled = false;
  double kernel_ms[RX_N_TIMED] = {};
  int64_t counters[3] = {};  // items, FP64 re-checks, re-check overflow (last run)
  int shape_class = 0;
  uint64_t config_gen = 0;  // Global::config_gen at build time
  // graph of the launch sequence (captured on the second identical run)
  RunKey last_key, graph_key;
  bool last_valid = false, graph_failed = false;
  // one-shot batches (the server's slots: every run carries new requests)
  // never capture a launch graph, whose capture would cost more than it saves
  bool one_shot = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int64_t graph_replays = 0;
  ~Batch();
};

// Fills b (re-using device buffers when large enough) from the requests.
int batch_build(Batch &b, const rx_request *req, int n_req, std::string &err);
// Chooses the work-item segment length and the persistent wavefronts per CU
// (0: automatic / autotuned) and sizes the work list.
void batch_plan(Batch &b, int seg_bins, int waves_per_cu);
int batch_upload(Batch &b);
int batch_run(Batch &b, int precision, std::string &err);
int batch_sync(Batch &b);
// Public status of a request from its device status.
int public_status(int dev_status);

// One resident xillver table: the host table, why a table named by
// RELXILL_TABLE_PATH could not be loaded (empty: loaded), its energy edges
// (elo, then the last ehi) and the device copies (renormalised spectra,
// edges, the Ecut / kTe axis, the axis ensure_ecut_within_boundarys reads, the
// inclinations).
struct XillSlot {
  XillTableHost t;
  bool loaded = false, dirty = true;
  std::string err;
  std::vector<double> ener;
  DevBuf spec, d_ener, ect, q, incl, lxi;
};

// global state
struct Global {
  std::recursive_mutex mu;
  bool device_ready = false;
  int device = 0;
  int n_cu = 256;
  RelTableHost rel;
  LpTableHost lp;
  bool rel_loaded = false, lp_loaded = false;
  // why a table named by RELXILL_TABLE_PATH could not be loaded (empty: loaded)
  std::string rel_err, lp_err, rrad_err;
  bool rel_dirty = true, lp_dirty = true;
  double rel_gmin_all = 0.0, rel_gmax_all = 1e30;  // g range reachable by any window
  DevBuf t_a, t_mu0, t_r, t_gmin, t_gmax, t_trff1, t_trff2, t_cosne1, t_cosne2;
  DevBuf l_a, l_h, l_rad, l_int, l_del, l_inc;
  // xillver tables, one per primary continuum (PRIM_*: xillver-a-Ec5.fits for
  // the cut-off power law, xillverCp_v3.4.fits for nthcomp, xillverNS-2.fits
  // for the blackbody, src/common.h:163-167;
  // loaded and uploaded on the first relxill request that needs one)
  XillSlot xill[N_PRIM];
  // returning-radiation table (loaded and uploaded on the first request with it)
  RradTableHost rrad;
  bool rrad_loaded = false, rrad_dirty = true;
  DevBuf r_rlo, r_rhi, r_tf, r_gmin, r_gmax, r_fg;
  // convolution grid
  int conv_n = 4096;
  double conv_emin = 0.00035, conv_emax = 2000.0;
  std::vector<double> conv_ener, conv_cfac, conv_cfinv, conv_inv_emid, conv_tw, conv_twp_fwd, conv_twp_inv;
  int conv_i1kev = 0;
  int conv_band_lo = 0, conv_band_hi = 0;
  bool conv_dirty = true;
  std::vector<double> coarse_ener, coarse_ln, conv_ln_emid;
  DevBuf c_coarse, c_coarse_ln, c_ln_emid;
  DevBuf c_ener, c_cfac, c_cfinv, c_inv_emid, c_tw, c_twp_fwd, c_twp_inv;
  int nr_fine = 1000;
  // reflection sets
  std::vector<DevReflSet> refl_sets;
  std::vector<DevBuf *> refl_bufs;
  DevBuf d_refl_sets;
  bool refl_dirty = false;
  bool profile = false;
  bool use_graphs = true;  // replay repeated launch sequences as hipGraphs
  // bumped by every change that invalidates built batches (tables, conv grid,
  // fine-grid size, clearing reflection sets); a batch refuses to run on a
  // newer configuration than it was built for
  uint64_t config_gen = 1;
  float guard = ROMB_GUARD;
  // autotune cache: shape class -> (segment bins, waves per CU)
  struct Tuned {
    int shape, seg_bins, waves_per_cu;
  };
  std::vector<Tuned> tune;
};

Global &G();
void ensure_tables_host();
int ensure_device(std::string &err);  // device + uploads (tables, conv grid, refl sets)
// rx_nthcomp: n_spec nthcomp spectra on one caller grid (host arrays)
int nthcomp_spectra(const double *ener, int n_flux, int n_spec, const double *kte, const double *gam,
                    const double *ener_shift, double *flux, std::string &err);
void ensure_xilltable_host(int prim);
void ensure_rradtable_host();
int ensure_rradtable_device(std::string &err);
// the returning-radiation table zones of an LP request (select_spinIndexForTable
// and allocate_radial_grid, src/Relreturn_Table.cpp:343-404); a status != 0 if
// the table does not cover the disk
int rrad_zones(const RradTableHost &t, ReqParams &p, int nr_fine);
int ensure_xilltable_device(int prim, std::string &err);
void fill_dev_views(BatchDev &bd);
void set_last(int status, const std::string &msg);
// the same with a static message (no allocation)
void set_last_static(int status, const char *msg);
// Looks up the device address of the allocation-free entry point's scratch
// array on `device` (current) once (noalloc.hip).
int noalloc_setup(int device, std::string &err);
int last_status();
const char *last_error();
bool phys_norm_env();

}  // namespace rx
