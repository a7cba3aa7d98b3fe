// noalloc.hip -- the allocation-free evaluation entry point.
//
// rx_eval_model_noalloc evaluates one line or convolution request on device
// data that the caller owns, with every work buffer of the kernels (radius
// records, work list, FP64 re-check queue, counters, profile)
// carved from one fixed device array, `scratch`, instead of memory the
// library allocates.  The array's device address is looked up once per
// device when the library initialises the device; host code only ever uses
// that address.  Every call on a GPU shares the array and leaves the last
// call's contents in it: whatever must start at zero is cleared by the
// kernels themselves (k_reset clears the counters, k_setup the profile bins
// that windows reach, k_finalize / k_conv write every output bin), and calls
// that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "engine.h"
#include "rx_math.h"

namespace rx {

namespace {

constexpr size_t kScratchBytes = RX_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

// the work buffers of rx_eval_model_noalloc (one copy per GPU)
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

void *g_scratch[kMaxDevices] = {};  // its device address, per device

// relconv's input check (src/LocalModel.cpp:84-86) on device data: an input
// spectrum whose sum is not positive fails the request with status 50.
constexpr int CHECK_THREADS = 256;
__global__ __launch_bounds__(CHECK_THREADS) void k_input_check(const double *flux, int n, ReqParams *req,
                                                               int *status) {
  __shared__ double s[CHECK_THREADS];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += CHECK_THREADS) v += flux[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = CHECK_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && !(s[0] > 0.0)) {
    req->pre_status = RX_ERR_INPUT;
    *status = RX_ERR_INPUT;
  }
}

// Bump allocation of the request's buffers inside the scratch array.
struct Carve {
  char *base;
  size_t used = 0;
  template <class T>
  T *take(size_t n) {
    used = (used + 255) & ~(size_t)255;
    T *p = reinterpret_cast<T *>(base + used);
    used += sizeof(T) * (n ? n : 1);
    return p;
  }
};

}  // namespace

int noalloc_setup(int device, std::string &err) {
  if (device < 0 || device >= kMaxDevices) {
    err = "device id out of range for the allocation-free entry point";
    return RX_ERR_DEVICE;
  }
  if (g_scratch[device]) return RX_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    err = "hipGetSymbolAddress of the scratch array failed";
    return RX_ERR_DEVICE;
  }
  g_scratch[device] = p;
  return RX_OK;
}

namespace {

int noalloc_eval(int model, const double *parameter, const double *d_energy, const double *d_flux_in,
                 double *d_flux_out, int n_flux, int precision, hipStream_t s, const char *&msg) {
  Global &g = G();
  if (!parameter || !d_energy || !d_flux_out || n_flux <= 0 || (precision != RX_FP64 && precision != RX_FP32)) {
    msg = "null pointer, empty grid or unknown precision";
    return RX_ERR_ARG;
  }
  const bool conv = (model == MOD_RELCONV || model == MOD_RELCONVLP);
  if (!(conv || model == MOD_RELLINE || model == MOD_RELLINELP)) {
    msg = "the allocation-free entry point evaluates relline, relline_lp, relconv and relconv_lp";
    return RX_ERR_ARG;
  }
  if (conv && !d_flux_in) {
    msg = "convolution request without input spectrum";
    return RX_ERR_ARG;
  }
  if (!g.device_ready || !g.rel_loaded || g.rel_dirty || g.conv_dirty ||
      (model != MOD_RELLINE && model != MOD_RELCONV && (!g.lp_loaded || g.lp_dirty))) {
    msg = "tables or convolution grid not resident: call rx_init() after changing them";
    return RX_ERR_ARG;
  }
  if (g.device < 0 || g.device >= kMaxDevices || !g_scratch[g.device]) {
    msg = "scratch array of the device not set up";
    return RX_ERR_DEVICE;
  }
  if (hipSetDevice(g.device) != hipSuccess) {
    msg = "hipSetDevice failed";
    return RX_ERR_DEVICE;
  }
  ReqParams p;
  std::string perr;
  int st = parse_request(model, parameter, phys_norm_env(), p, perr);
  if (st == RX_ERR_ARG) {
    msg = "unknown model id";
    return st;
  }
  p.n_in = n_flux;
  p.ener_off = 0;
  p.flux_off = 0;
  p.refl_set = -1;
  p.n_zones = 1;
  p.n_e = conv ? g.conv_n : n_flux;
  p.pgrid_off = conv ? -1 : 0;
  p.prof_off = conv ? 0 : -1;
  p.dist_off = 0;
  p.node_off = -1;
  p.zone_off = 0;
  p.xspec_off = 0;
  if (st == RX_OK && p.return_rad != 0) {
    if (!g.rrad_loaded || g.rrad_dirty) {
      msg = "returning-radiation table not resident: evaluate one request with it through rx_eval_model first";
      return RX_ERR_ARG;
    }
    st = rrad_zones(g.rrad, p, g.nr_fine);
  }
  p.need_cosne = (p.limb != 0) ? 1 : 0;
  p.pre_status = st;

  // work shape: the autotuned one of the request's shape class, else the default
  int lg = 0;
  while ((1LL << lg) < p.n_e) lg++;
  const int shape = lg * 2 + (conv ? 1 : 0);
  int seg = 0, wpc = 0;
  for (const auto &t : g.tune)
    if (t.shape == shape) {
      seg = t.seg_bins;
      wpc = t.waves_per_cu;
    }
  if (seg <= 0) seg = 64;
  if (wpc <= 0) wpc = 8;
  const int n_rc = (g.nr_fine + RAD_CHUNK - 1) / RAD_CHUNK;
  const int item_cap = n_rc * ((p.n_e + seg - 1) / seg);
  const int fix_cap = std::max(4096, p.n_e * 4);

  BatchDev bd;
  memset(&bd, 0, sizeof(bd));
  fill_dev_views(bd);
  Carve c{static_cast<char *>(g_scratch[g.device])};
  bd.n_req = 1;
  bd.nr_fine = g.nr_fine;
  bd.seg_bins = seg;
  bd.node_stride = 2 * g.rel.n_r * g.rel.n_g * 2;
  bd.req = c.take<ReqParams>(1);
  bd.ener_in = d_energy;
  bd.flux_in = conv ? d_flux_in : nullptr;
  bd.out = d_flux_out;
  bd.pgrid = conv ? nullptr : c.take<double>((size_t)n_flux + 1);
  bd.node = nullptr;  // line and convolution requests read no node rows
  bd.rad = c.take<RadRec>((size_t)g.nr_fine);
  bd.rnode = c.take<double>((size_t)RAD_NODE_ARRAYS * g.rel.n_r);
  bd.rnorm = c.take<ReqNorm>(1);
  bd.prof = conv ? c.take<double>((size_t)g.conv_n) : nullptr;
  bd.dist = c.take<double>(1);
  bd.refl = c.take<LpRefl>(1);
  bd.status = c.take<int>(1);
  bd.items = c.take<WorkItem>((size_t)item_cap);
  bd.item_cap = item_cap;
  bd.ctr = c.take<int>(N_COUNTERS);
  bd.reach = c.take<BinRange>(1);
  bd.fix = c.take<FixEntry>((size_t)fix_cap);
  bd.fix_cap = fix_cap;
  bd.guard = g.guard;
  bd.xspec = nullptr;
  bd.zone = c.take<ZoneRec>(1);
  XillWork *d_work = c.take<XillWork>(1);
  if (c.used > kScratchBytes) {
    msg = "request needs more work memory than the allocation-free entry point holds "
          "(RX_NOALLOC_SCRATCH_BYTES): fewer bins or smaller tables";
    return RX_ERR_ARG;
  }
  const XillWork w0{0, 0};
  int e = (int)hipMemcpyAsync(bd.req, &p, sizeof(p), hipMemcpyHostToDevice, s);
  e |= (int)hipMemcpyAsync(bd.status, &p.pre_status, sizeof(int), hipMemcpyHostToDevice, s);
  e |= (int)hipMemcpyAsync(d_work, &w0, sizeof(w0), hipMemcpyHostToDevice, s);
  const bool cos_rows = p.limb != 0;
  const int n_blocks = std::min(g.n_cu * wpc, std::max(1, item_cap));  // no more wavefronts than items
  int out[2] = {0, 0};  // status, FP64 re-check queue overflow
  for (int pass = 0; pass < 2 && e == 0; pass++) {
    bd.mode = (pass == 0 && precision == RX_FP32) ? 1 : 0;
    if (conv && p.pre_status == 0) {
      hipLaunchKernelGGL(k_input_check, dim3(1), dim3(CHECK_THREADS), 0, s, d_flux_in, n_flux, bd.req,
                         bd.status);
      e |= (int)hipGetLastError();
    }
    e |= launch_setup(bd, s);
    e |= launch_integrate(bd, n_blocks, cos_rows, s);
    e |= launch_fixup(bd, cos_rows, s);
    e |= launch_finalize(bd, s);
    if (conv) e |= launch_conv(bd, d_work, 1, s);
    e |= (int)hipMemcpyAsync(&out[0], bd.status, sizeof(int), hipMemcpyDeviceToHost, s);
    e |= (int)hipMemcpyAsync(&out[1], bd.ctr + CTR_FIX_OVF, sizeof(int), hipMemcpyDeviceToHost, s);
    e |= (int)hipStreamSynchronize(s);
    // an overflowing FP32 re-check queue: the pass is redone in FP64
    if (!(bd.mode == 1 && out[1] != 0)) break;
  }
  if (e) {
    (void)hipGetLastError();
    msg = "device error in the allocation-free evaluation";
    return RX_ERR_DEVICE;
  }
  if (out[0] != 0) {
    msg = "model evaluation failed";
    return public_status(out[0]);
  }
  msg = "";
  return RX_OK;
}

}  // namespace
}  // namespace rx

extern "C" int rx_eval_model_noalloc(int model, const double *parameter, const double *d_energy,
                                     const double *d_flux_in, double *d_flux_out, int n_flux, int precision,
                                     void *stream) {
  using namespace rx;
  Global &g = G();
  std::lock_guard<std::recursive_mutex> lk(g.mu);
  const char *msg = "";
  int st;
  try {
    st = noalloc_eval(model, parameter, d_energy, d_flux_in, d_flux_out, n_flux, precision,
                      static_cast<hipStream_t>(stream), msg);
  } catch (...) {
    st = RX_ERR_DEVICE;
    msg = "unexpected exception";
  }
  set_last_static(st, msg);
  return st;
}
