// rx_engine.h -- device-side data layout of a batch of spectra and the kernel
// launch interface (internal to librelxill_gpu).
//
// HBM layout of one batch (all arrays flat, offsets per request in ReqParams):
//   req[n]              processed model parameters (host mirror of
//                       get_rel_params, src/ModelDefinition.cpp:278-334)
//   ener_in             caller energy grids, n_in+1 edges each
//   flux_in             caller input spectra (convolution models)
//   out                 caller-grid outputs, n_in each (line models accumulate
//                       their profile here directly)
//   pgrid               profile grids (line models: the caller grid shifted
//                       to line-energy units; conv models use the shared
//                       convolution grid instead)
//   node[n]             (a,mu0)-interpolated table node, FP64:
//                       trff[n_r][n_g][2] followed by cosne[n_r][n_g][2]
//   rad[n][nr_fine]     per fine radius record (RadRec, 16 bytes)
//   rnode[n][3][n_r]    per request: the (a, mu0)-interpolated node radii,
//                       g_min and g_max that the records' geometry is rebuilt from
//   prof                per request profile on the conv grid (conv models)
//   items               work list written by k_setup: (request, chunk of 64
//                       fine radii, segment of at most seg_bins bins)
//   dist[n][n_incl]     emission-angle distribution (relxill-style requests)
//   xspec               relxill requests: the angle-weighted xillver spectrum
//                       (n_ener of the xillver table each), written by k_xill
#pragma once
#include <stdint.h>

namespace rx {

constexpr int MAX_NR_TAB = 256;   // table radii held in LDS by the setup kernel
constexpr int MAX_NG = 64;        // g* points (trff values per radius: 2*ng)
constexpr int MAX_LP_R = 256;     // lamp-post table radii
constexpr int MAX_INCL = 64;      // emission-angle bins
constexpr int MAX_FFT = 8192;     // convolution grid size limit
constexpr int LP_ZONES = 10;      // radial zones of relxilllp (N_ZONES, src/common.h:134)
constexpr int MAX_ZONES = 25;     // with an ionisation gradient (N_ZONES_IONGRAD, src/common.h:135)
constexpr int MAX_RRC_ENER = 16384;  // xillver bins held in LDS by the returning-radiation corrections

enum ModelKind : int {
  MOD_RELLINE = 1,
  MOD_RELLINELP = 2,
  MOD_RELCONV = 11,
  MOD_RELCONVLP = 12,
  MOD_RELXILL = -1,          // relxill: broken power law, one zone, xillver table
  MOD_RELXILLLP = -2,        // relxilllp: lamp post, LP_ZONES radial zones (MAX_ZONES with a gradient)
  MOD_RELXILL_CONV = -1001,  // relxill-style conv of a resident reflection set
  // the nthcomp-primary models (relxillCp, relxilllpCp); their requests run as
  // MOD_RELXILL / MOD_RELXILLLP with ReqParams.prim = PRIM_NTHCOMP
  MOD_RELXILLCP = -1010,
  MOD_RELXILLLPCP = -1011,
  // the blackbody-primary relxill model (relxillNS), run as MOD_RELXILL with
  // ReqParams.prim = PRIM_BBODY
  MOD_RELXILLNS = -1012,
  // the non-relativistic xillver models (the reference's MOD_TYPE_XILLVER,
  // _XILLVER_NTHCOMP, _XILLVERNS, src/common.h:91-111); their requests run as
  // MOD_XILLVER with ReqParams.prim naming the primary
  MOD_XILLVER = 0,
  MOD_XILLVERCP = 100,
  MOD_XILLVERNS = -101,
};

enum EmisKind : int { EMIS_BKN = 1, EMIS_LP = 2 };

struct ReqParams {
  int model;
  int emis_type;
  int limb;
  int do_renorm;
  int n_e;        // bins of the profile grid
  int n_in;       // bins of the caller grid
  int need_cosne; // limb != 0 or angle distribution requested
  int n_incl;     // >0: compute the angle distribution with n_incl bins
  double a, incl, emis1, emis2, rbr, rin, rout, lineE, z, height, gamma, beta;
  double norm;    // renormalisation target (1, or 0.5 cos i for relxill BKN)
  double rms;     // ISCO of a (host kerr_rms: Rin was reset to exactly this value)
  int64_t ener_off;   // caller grid edges in ener_in
  int64_t flux_off;   // caller input / output spectrum offset (n_in)
  int64_t pgrid_off;  // profile grid edges in pgrid (-1: shared conv grid)
  int64_t prof_off;   // profile (n_e) for conv models, -1 otherwise
  int refl_set;       // relxill-style: which resident reflection set
  int pre_status;     // host-side parameter status; nonzero requests are skipped
  // relxill: xillver corner (lower indices of the four interpolated table
  // parameters, in table order) and interpolation factors, the primary
  // continuum and the offset of the request's xillver spectrum in xspec
  int xind[4];
  double xfac[4];
  double xgam, xect, refl_frac;
  int64_t xspec_off;  // n_zones spectra of the table's n_ener
  // radial zones (relxill models): edges (RadialGrid, src/IonGradient.cpp:457-489);
  // dist holds n_zones x n_incl values from dist_off, prof n_zones x n_e
  int n_zones;
  int xboost;         // relxilllp switch_reflfrac_boost
  double zgrid[MAX_ZONES + 1];
  double g_so;        // energy shift primary source -> observer (1 without a lamp post)
  double xect_src;    // Ecut in the source frame
  int64_t dist_off;
  int64_t zone_off;   // ZoneRec per zone
  // the request's (a, mu0)-interpolated node rows in node (n_incl > 0 only:
  // the angle distribution reads them; -1 otherwise)
  int64_t node_off;
  // primary continuum of the relxill models (PRIM_*; it also selects the
  // xillver table, b.xt[prim]).  nthcomp: n_nth solves (the source's kTe, then
  // the kTe every lamp-post zone sees) with their results at nth[nth_off],
  // the source's solution in slot nth_sol.  xdens_div: the density division
  // of the table rows (renorm_xill_spec, src/xilltable.c:478-486) for tables
  // without a density axis, 0 when logN is 15.
  int prim, n_nth, nth_sol, pad_prim;
  int64_t nth_off;
  double xdens_div;
  double xktbb;  // blackbody primary: kT_bb (keV)
  // the xillver models: the inclination's table bracket and float factor
  // (the fifth interpolated parameter), a two-bin angle distribution for k_xill
  int xinc_ind, pad_xinc;
  double xinc_fac;
  // a six-parameter table: the bracket and factor of its first axis, which is
  // interpolated outside the other five (0 and 0 for five parameters)
  int xo_ind, pad_xo;
  double xo_fac;
  // relxilllpCp's ionisation gradient (iongrad_type, src/IonGradient.cpp:
  // 340-395): 0 constant, 1 the power law of calc_ion_grad_pl (:174-182) in
  // the zones' radius with index ion_index from the input logXi xlxi0
  int ion_grad, pad_ion;
  double ion_index, xlxi0;
  // returning radiation (lamp post): 1 add, -1 or 2 replace the direct
  // emissivity; the table spin, its first zone and the number of zones used
  int return_rad, rr_k, rr_klo, rr_n;
  // relxilllp above spin 0: the zones' xillver correction factors of the
  // returning radiation (flux, g shift per zone) at rrcorr[rr_corr_off]
  int rr_corr;
  int64_t rr_corr_off;
};

// Per zone of a relxill request (IonGradient, src/IonGradient.cpp:182-209):
// the source -> disk energy shift of the zone centre, the Ecut seen there,
// its xillver axis bracket and factor, and the normalisation change of the
// incident spectrum (src/Xillspec.cpp:408-435).
struct ZoneRec {
  double rmean, del_emit, shift, ncf, ect_fac, lxi_fac;
  int ect_ind, lxi_ind;  // the zone's brackets on the Ecut and logXi axes (with ect_fac, lxi_fac)
};

// A k_conv unit: a request and one of its zones.  k_xill units use the same
// record with zone 0: one-zone requests, or all zones of a lamp-post request.
struct XillWork {
  int req, zone;
};

// One fine radius as k_setup stores it (16 bytes).  The bin integral at this
// radius is
//   flux[j] += integ_bin(E_j, E_j+1) * coef,   coef = emis*weight/(gmax-gmin)
// The radius itself, its lerp factor between the node radii and its g window
// are not stored: rad_geo rebuilds them from the radius index, the bracket K
// and the request's node arrays (RAD_NODE_ARRAYS x n_r doubles per request:
// node radii, g_min, g_max of the nodes), bit for bit as k_setup computed them.
struct RadRec {
  double coef;  // emis / (gmax - gmin) * weight, times ReqNorm.fac
  uint32_t lo;  // first bin of the window (bits 0-23) | K << 24
  uint32_t hi;  // last bin of the window, inclusive (bits 0-23) | zone << 24
};
constexpr int RAD_BIN_BITS = 24;  // bins of a profile grid: fewer than 2^24
constexpr uint32_t RAD_BIN_MASK = (1u << RAD_BIN_BITS) - 1;
constexpr int RAD_NODE_ARRAYS = 3;
// (ielo > iehi: the radius does not contribute)
__host__ __device__ inline int rad_ielo(const RadRec &r) { return (int)(r.lo & RAD_BIN_MASK); }
__host__ __device__ inline int rad_iehi(const RadRec &r) { return (int)(r.hi & RAD_BIN_MASK); }
__host__ __device__ inline int rad_K(const RadRec &r) { return (int)(r.lo >> RAD_BIN_BITS); }
__host__ __device__ inline int rad_zone(const RadRec &r) { return (int)(r.hi >> RAD_BIN_BITS); }
__host__ __device__ inline RadRec rad_pack(double coef, int K, int zone, int ielo, int iehi) {
  RadRec r;
  r.coef = coef;
  r.lo = (uint32_t)ielo | ((uint32_t)K << RAD_BIN_BITS);
  r.hi = (uint32_t)iehi | ((uint32_t)zone << RAD_BIN_BITS);
  return r;
}
static_assert(MAX_NR_TAB <= 256 && MAX_ZONES <= 256, "K and the zone take 8 bits of a radius record");

// Per-request factors of the radius weights.  RadRec.coef holds the weight
// without the emissivity normalisation; fac is that normalisation (1/area for
// the broken power law, 1 for the lamp post).  The FP32 path sums the 64
// per-radius values of a bin in FP32; cmul = fac 2^k brings the request's
// largest weight near 1 for that (far from FP32 under- and overflow) and
// cinv = 2^-k undoes it exactly.
struct ReqNorm {
  double fac, cmul, cinv;
  // the (a, mu0) corners of the request's table interpolation (interpol_a_mu0,
  // src/Relprofile.cpp:39-80): spin index ia, mu0 index im and their factors;
  // k_integrate and k_fixup build the node rows they need from these
  double fa, fm;
  int ia, im;
  // 1/sqrt(Rout) and 1/sqrt(Rin) of the fine radial grid (fine_radius_at)
  double r1, r2;
};

struct LpRefl {
  double refl_frac, f_bh, f_ad, f_inf, f_inf_rest;
};

struct DevRelTable {
  int n_a, n_mu0, n_r, n_g;
  const float *a, *mu0, *r, *gmin, *gmax, *trff1, *trff2, *cosne1, *cosne2;
};

// Returning-radiation table resident on the device (flat, spin-major).
constexpr int MAX_RR = 64;  // table zones a request may use
struct DevRradTable {
  int n_spin, n_rad, n_g;
  const double *rlo, *rhi, *tf_r, *gmin, *gmax, *frac_g;
};

struct DevLpTable {
  int n_a, n_h, n_r;
  const float *a, *h, *rad, *intens, *del, *del_inc;
};

// Reflection set resident on the device for relxill-style requests: the
// per-angle spectra rebinned to the convolution grid, their band sums and
// the real FFT of (spectrum * E/dE) on the convolution grid.
struct DevReflSet {
  int n_incl;
  const double *spec;    // [n_incl][n_conv]   rebinned, photons/bin
  const double *bandsum; // [n_incl]
  const double *fft;     // [n_incl][n_conv/2+1][2]
};

// Resident xillver table: the spectra divided by 10^logXi at load time,
// row ((((i0 n1 + i1) n2 + i2) n3 + i3) n_incl + ia) at spec[row * n_ener].
// primary continua of the relxill models (PRIM_SPEC_*, src/common.h:57-59)
enum PrimType : int { PRIM_CUTOFFPL = 0, PRIM_NTHCOMP = 1, PRIM_BBODY = 2 };
constexpr int N_PRIM = 3;

// Normalisation results of one nthcomp solve: nf = 1 / calcNormWrtXillverTableSpec
// of the spectrum on the 500-bin 0.1-1000 keV grid (src/Xillspec.cpp:179-205)
// and pe, the energy flux of the normalised spectrum in 0.1-1000 keV
// (get_energy_flux_band, :132-144).
struct NthRes {
  double nf, pe;
};

struct DevXillTable {
  int n_ener, n_incl;
  int n_val[4];         // values of the four (inner) interpolated parameters
  // a six-parameter table: n_out values of its first axis, one step of which
  // moves a row by outer_rows rows (n_out 1, outer_rows 0: five parameters)
  int n_out;
  unsigned outer_rows;
  const float *spec;
  const double *ener;   // n_ener + 1 edges (elo, then the last ehi)
  int ect_dim;          // table axis of Ecut (-1: none)
  int ect_q;            // the axis ensure_ecut_within_boundarys reads (src/xilltable.c:1058-1066)
  const float *ect_vals;  // values of axis ect_dim
  const float *q_vals;    // values of axis ect_q
  int n_ect, n_q;
  const float *incl_vals;  // the inclination axis (degrees)
  int lxi_dim;             // inner axis of logXi (-1: none)
  int n_lxi;
  const float *lxi_vals;   // its values (the zones of an ionisation gradient bracket their own)
};

constexpr int N_ENER_COARSE = 500;  // primary-continuum normalisation grid (src/Xillspec.h:32)

struct DevConvGrid {
  int n;               // bins (power of two)
  int i1kev;           // binary_search(ener, n+1, 1.0)
  int band_lo, band_hi;  // bins with E_lo >= 0.01 and E_hi < 1000 keV: [band_lo, band_hi)
  const double *ener;  // n+1 edges
  const double *cfac;  // E/dE conversion, n
  const double *cfinv; // dE/E, n
  const double *inv_emid;  // 1/E_mid, n
  const double *tw;    // twiddles exp(-2 pi i k/n), k < n/2, interleaved re/im
  const double *twp_fwd;  // per-pass twiddles of the n-point FFT ([r][k] per pass)
  const double *twp_inv;  // per-pass twiddles of the n/2-point FFT
  const double *coarse;   // N_ENER_COARSE + 1 log edges over 0.1-1000 keV
  const double *coarse_ln;  // ln of the coarse bin centres
  const double *ln_emid;    // ln of the bin centres, n
};

// Fine radii are integrated in chunks of one wavefront: lane l of chunk c
// owns fine radius 64 c + l.
constexpr int RAD_CHUNK = 64;
constexpr int MAX_NR_FINE = 64 * RAD_CHUNK;

// k_integrate parks the per-radius values of SUM_BINS consecutive bins in LDS
// ([bin][lane], row stride SUM_STRIDE: lane (c, p) of the turn-round reads
// values SUM_BINS p .. SUM_BINS p + SUM_BINS - 1 of bin c, and the odd stride
// spreads the bins over the banks) and then sums each bin with 64 / SUM_BINS
// lanes (SUM_BINS radii each, then a DPP tree); SUM_BINS is 8 or 4.
constexpr int SUM_BINS = 8;
constexpr int SUM_STRIDE = 65;
static_assert(SUM_BINS == 8 || SUM_BINS == 4, "turn-round DPP tree covers 8 or 16 lanes");

// One unit of integration work: bins [j0, j1) of one request for the fine
// radii of one chunk.  k_setup emits the segments of each chunk's energy
// window (union over its radii), k_integrate pulls them from a shared counter.
struct WorkItem {
  int req, rc, j0, j1;
};

struct BinRange {
  int lo, hi;
};

// FP32 bins whose Romberg stop test fell in the guard band: recomputed in FP64
// by k_fixup, which adds (FP64 value - FP32 value) to the profile.
struct FixEntry {
  int req, rad, bin, pad;
  double v32;  // contribution already accumulated (float value, widened)
};

// k_setup status of a request whose work items did not fit the work list.
constexpr int ST_WORKLIST = 60;

// k_integrate splits the work list into N_XQ contiguous ranges, one per XCD
// of the MI355X (workgroups are dispatched to the XCDs round-robin), each
// handed out by its own counter; a wavefront takes items of its own range
// first, so the items of one request (which share node rows) are integrated
// from one XCD's L2, and moves on to the next range when its own is empty.
constexpr int N_XQ = 8;
constexpr int XQ_STRIDE = 32;  // ints between the range counters (one 128-byte line each)

// Device counters of one batch run (cleared by k_reset before k_setup).
enum CounterSlot : int {
  CTR_FIX = 0,        // FixEntry entries written
  CTR_FIX_OVF = 1,    // FixEntry queue overflowed
  CTR_ITEMS = 2,      // long work items written (from the front of the list)
  CTR_ITEMS_OVF = 3,  // work list overflowed
  CTR_ITEMS_S = 4,    // short work items written (from the back of the list)
  CTR_XQ = XQ_STRIDE,  // N_XQ range counters, XQ_STRIDE apart
  N_COUNTERS = CTR_XQ + N_XQ * XQ_STRIDE
};

struct BatchDev {
  int n_req;
  int nr_fine;
  int seg_bins;     // bins per work item (at most)
  int node_stride;  // doubles per request in node (2 * n_r * n_g * 2)
  int max_zones;    // most zones of a request (k_setup's zone arrays: LP_ZONES or MAX_ZONES)
  ReqParams *req;
  const double *ener_in;
  const double *flux_in;
  double *out;
  double *pgrid;
  double *node;
  RadRec *rad;
  double *rnode;  // per request: RAD_NODE_ARRAYS * rt.n_r doubles (node radii, g_min, g_max)
  ReqNorm *rnorm;
  double *prof;
  double *dist;
  LpRefl *refl;
  int *status;
  WorkItem *items;
  int item_cap;
  BinRange *reach;    // per request: bins [lo, hi] that some window reaches (lo > hi: none)
  int *ctr;           // CounterSlot
  FixEntry *fix;
  int fix_cap;
  int mode;           // 0 FP64, 1 FP32
  float guard;        // FP32 Romberg guard band (relative)
  DevRelTable rt;
  DevLpTable lp;
  DevConvGrid cg;
  const DevReflSet *refl_sets;
  DevXillTable xt[N_PRIM];  // one xillver table per primary continuum
  double *xspec;
  ZoneRec *zone;
  DevRradTable rr;
  double *rrcorr;     // relxilllp returning radiation: 2 factors per zone
  // nthcomp primary: per solve its normalisation results, per request the
  // source's solution (x: NTH_MAX_POINTS + 1, E F_E: NTH_MAX_POINTS, jmax)
  NthRes *nth;
  double *nth_x, *nth_s;
  int *nth_n;
};

// ---- launchers (kernels.hip) ----
int launch_setup(const BatchDev &b, void *stream);
int launch_integrate(const BatchDev &b, int n_blocks, bool limb, void *stream);
int launch_fixup(const BatchDev &b, bool cos_rows, void *stream);
int launch_finalize(const BatchDev &b, void *stream);
// conv_work: (request, zone) pairs; relxill requests have one per zone
int launch_conv(const BatchDev &b, const XillWork *conv_work, int n_conv_work, void *stream);
// work: n_single one-zone units, then n_multi units of several zones
int launch_xill(const BatchDev &b, const XillWork *work, int n_single, int n_multi, void *stream);
// relxilllp returning radiation above spin 0: the zones' xillver correction
// factors, one workgroup per (request, zone) of work (before launch_setup)
int launch_rrad_corr(const BatchDev &b, const XillWork *work, int n_work, void *stream);
// relxill requests: the primary continuum added to the output (after launch_conv)
int launch_primary(const BatchDev &b, const int *reqs, int n_reqs, void *stream);
int launch_copy(const void *src, void *dst, size_t bytes, void *stream);

// ---- nthcomp primary continuum (nthcomp.hip) ----
// photon energies of one Kompaneets solve (the reference's arrays of 900)
constexpr int NTH_MAX_POINTS = 900;
struct NthJob {
  double kte, gam, ktbb;  // keV, photon index, seed temperature (keV)
};
// one solve per job: x (NTH_MAX_POINTS + 1 per job, m_e c^2), E F_E
// (NTH_MAX_POINTS per job), the number of points
int launch_nthcomp(const NthJob *jobs, int n_jobs, double *sol_x, double *sol_s, int *sol_n, void *stream);
// relxill models with the nthcomp primary: one solve per (request, j) of
// work (j = 0 the source, j = 1 + z zone z of a lamp post), first in a run
int launch_nth(const BatchDev &b, const XillWork *work, int n_work, void *stream);
int launch_refl_prepare(const DevConvGrid &cg, const double *refl_ener, const double *refl,
                        int n_refl, int n_incl, double *spec, double *bandsum, double *fft,
                        void *stream);

}  // namespace rx
