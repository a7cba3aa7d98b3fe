// server.cpp -- continuous-batching serving loop.
//
// Requests are copied at submit time into a pending queue.  A scheduler thread
// keeps up to n_streams batches on the device at once, each on its own HIP
// stream: whenever a slot is free and requests are pending, it admits pending
// requests in arrival order while their device working set fits the HBM
// budget (minus what the running batches hold) and the in-flight cap, builds
// a batch (without holding the queue lock, so submits and polls go on) and
// launches it, followed on the same stream by the copies of its outputs,
// statuses and counters into the slot's pinned host buffers.  While every
// slot is busy, arrivals queue up and the next free slot takes all of them:
// batches grow with the load.  A finished batch (hipStreamQuery) has its
// results on the host already; its memory returns to the budget and its
// requests are marked done with their latency (submit -> completion seen).
// Tables and the convolution grid stay resident for the server's lifetime.
//
// Each request's work items depend only on the request itself and the
// batch's segment length; per-bin sums over radius chunks are added
// atomically in FP64, so a request's result equals its solo run up to FP64
// rounding of that order.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>
#include <cmath>
#include <vector>

#include "engine.h"

using namespace rx;

namespace {

using Clock = std::chrono::steady_clock;

// work-list sizing bound of the admission estimate (the smallest segment
// length the autotuner may choose)
constexpr int SERVER_MIN_SEG_BINS = 32;

struct Pending {
  int64_t ticket;
  int model;
  int refl_set;
  std::vector<double> energy, params, flux_in;
  Clock::time_point t_submit;
  int64_t bytes;
};

struct Result {
  bool done = false;
  int status = 0;
  double latency_ms = 0;
  std::vector<double> flux;
};

// Pinned host memory that a batch's outputs are copied into on its stream.
struct PinnedBuf {
  void *p = nullptr;
  size_t bytes = 0;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  bool ensure(size_t n) {
    if (p && bytes >= n) return true;
    if (p) n = std::max(n, bytes + bytes / 2);  // grows rarely (hipHostFree synchronises)
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    if (hipHostMalloc(&p, n, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return false;
    }
    bytes = n;
    return true;
  }
  template <class T>
  T *as() const { return reinterpret_cast<T *>(p); }
};

struct Slot {
  Batch batch;
  bool busy = false;
  std::vector<Pending> reqs;
  int64_t bytes = 0;
  PinnedBuf h_out, h_status, h_ctr;
};

int64_t request_bytes(const Pending &p) {
  Global &g = G();
  const bool lpx = (p.model == MOD_RELXILLLP || p.model == MOD_RELXILLLPCP);
  const bool xill = (p.model == MOD_XILLVER || p.model == MOD_XILLVERCP || p.model == MOD_XILLVERNS);
  const bool relxill = (p.model == MOD_RELXILL || p.model == MOD_RELXILLCP || p.model == MOD_RELXILLNS || lpx || xill);
  const bool conv = (p.model == MOD_RELCONV || p.model == MOD_RELCONVLP || p.model == MOD_RELXILL_CONV || relxill);
  // zones: LP_ZONES for the lamp post, MAX_ZONES for relxilllpCp with the
  // power-law ionisation gradient (iongrad_type, parameter 15)
  const bool grad = p.model == MOD_RELXILLLPCP && p.params.size() > 14 && std::lround(p.params[14]) == 1;
  const int64_t nz = grad ? MAX_ZONES : (lpx ? LP_ZONES : 1);
  const int64_t n_in = (int64_t)p.energy.size() - 1;
  const int64_t n_e = conv ? g.conv_n : n_in;
  const int64_t n_rc = (g.nr_fine + RAD_CHUNK - 1) / RAD_CHUNK;
  int64_t b = sizeof(ReqParams) + 8 * (2 * (n_in + 1) + 2 * n_in);
  b += 8 * (n_e + 1) + 8 * n_e * nz;
  b += (int64_t)sizeof(WorkItem) * (n_rc + nz - 1) * ((n_e + SERVER_MIN_SEG_BINS - 1) / SERVER_MIN_SEG_BINS);
  b += (int64_t)sizeof(FixEntry) * (n_e / 8 + 1);
  b += (int64_t)sizeof(RadRec) * g.nr_fine + 8 * RAD_NODE_ARRAYS * (int64_t)g.rel.n_r;
  b += 8 * 2 * (int64_t)g.rel.n_r * g.rel.n_g * 2;
  b += 8 * MAX_INCL * nz + sizeof(LpRefl) + sizeof(ReqNorm) + 64;
  if (relxill) {
    int64_t ne = 0;
    for (int pr = 0; pr < N_PRIM; pr++) ne = std::max<int64_t>(ne, g.xill[pr].t.n_ener);
    b += (8 * ne + (int64_t)(2 * sizeof(XillWork) + sizeof(ZoneRec)) + 16) * nz;
  }
  if (p.model == MOD_RELXILLCP || p.model == MOD_RELXILLLPCP || p.model == MOD_XILLVERCP)  // nthcomp solves
    b += (int64_t)(sizeof(XillWork) + sizeof(NthRes)) * (nz + 1) + 8 * (2 * NTH_MAX_POINTS + 1) + 4;
  return b;
}

}  // namespace

struct rx_server {
  int max_inflight, n_streams, precision;
  int64_t budget;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Pending> pending;
  std::unordered_map<int64_t, Result> results;
  std::vector<std::unique_ptr<Slot>> slots;
  std::atomic<bool> stop{false};
  std::thread th;
  int64_t next_ticket = 0;
  int64_t submitted = 0, completed = 0, batches = 0, max_concurrent = 0;
  int64_t inflight_bytes = 0;
  int inflight_reqs = 0;
  int device = 0;

  void loop();
  int launch(Slot &s, std::vector<Pending> &take, std::string &err);
  void launch_ready(std::unique_lock<std::mutex> &lk);
  bool reap(std::unique_lock<std::mutex> &lk);
};

// Builds, launches and queues the output copies of one batch on the slot's
// stream (called without the queue lock: only the scheduler thread touches
// the slots).
int rx_server::launch(Slot &s, std::vector<Pending> &take, std::string &err) {
  std::vector<rx_request> rq(take.size());
  for (size_t i = 0; i < take.size(); i++) {
    rq[i].model = take[i].model;
    rq[i].n_flux = (int)take[i].energy.size() - 1;
    rq[i].energy = take[i].energy.data();
    rq[i].parameter = take[i].params.data();
    rq[i].flux_in = take[i].flux_in.empty() ? nullptr : take[i].flux_in.data();
    rq[i].refl_set = take[i].refl_set;
  }
  std::lock_guard<std::recursive_mutex> lk(G().mu);
  int st = ensure_device(err);
  if (st != RX_OK) return st;
  // the autotuned work shape of the batch's shape class (else the default)
  s.batch.seg_forced = 0;
  s.batch.waves_forced = 0;
  s.batch.one_shot = true;
  st = batch_build(s.batch, rq.data(), (int)rq.size(), err);
  if (st != RX_OK) return st;
  if (!s.h_out.ensure(sizeof(double) * std::max<int64_t>(1, s.batch.tot_flux)) ||
      !s.h_status.ensure(sizeof(int) * s.batch.n_req) || !s.h_ctr.ensure(sizeof(int) * N_COUNTERS)) {
    err = "pinned host allocation failed";
    return RX_ERR_NOMEM;
  }
  st = batch_run(s.batch, precision, err);
  if (st != RX_OK) return st;
  hipStream_t q = s.batch.stream;
  int e = (int)hipMemcpyAsync(s.h_out.p, s.batch.d_out.p, sizeof(double) * s.batch.tot_flux, hipMemcpyDeviceToHost, q);
  e |= (int)hipMemcpyAsync(s.h_status.p, s.batch.d_status.p, sizeof(int) * s.batch.n_req, hipMemcpyDeviceToHost, q);
  e |= (int)hipMemcpyAsync(s.h_ctr.p, s.batch.d_ctr.p, sizeof(int) * N_COUNTERS, hipMemcpyDeviceToHost, q);
  if (e) {
    err = "output copy failed";
    return RX_ERR_DEVICE;
  }
  return RX_OK;
}

void rx_server::launch_ready(std::unique_lock<std::mutex> &lk) {
  // called with mu held; released while a batch is built and launched
  for (auto &sp : slots) {
    Slot &s = *sp;
    if (s.busy || pending.empty()) continue;
    std::vector<Pending> take;
    int64_t bytes = 0;
    while (!pending.empty()) {
      Pending &p = pending.front();
      if (inflight_reqs + (int)take.size() >= max_inflight) break;
      if (inflight_bytes + bytes + p.bytes > budget && !(take.empty() && inflight_reqs == 0)) break;
      bytes += p.bytes;
      take.push_back(std::move(p));
      pending.pop_front();
    }
    if (take.empty()) break;
    s.busy = true;  // reserved: its budget share is counted below
    inflight_bytes += bytes;
    inflight_reqs += (int)take.size();
    lk.unlock();
    std::string err;
    const int st = launch(s, take, err);
    lk.lock();
    if (st != RX_OK) {
      const auto now = Clock::now();
      for (auto &p : take) {
        Result &r = results[p.ticket];
        r.done = true;
        r.status = st;
        r.flux.assign(p.energy.size() - 1, 0.0);
        r.latency_ms = std::chrono::duration<double, std::milli>(now - p.t_submit).count();
        completed++;
      }
      s.busy = false;
      inflight_bytes -= bytes;
      inflight_reqs -= (int)take.size();
      cv.notify_all();
      continue;
    }
    s.reqs = std::move(take);
    s.bytes = bytes;
    batches++;
    int64_t conc = 0;
    for (auto &o : slots) conc += o->busy ? 1 : 0;
    if (conc > max_concurrent) max_concurrent = conc;
  }
}

bool rx_server::reap(std::unique_lock<std::mutex> &lk) {
  // called with mu held
  bool any = false;
  for (auto &sp : slots) {
    Slot &s = *sp;
    if (!s.busy || s.reqs.empty()) continue;
    hipError_t q = hipStreamQuery(s.batch.stream);
    if (q == hipErrorNotReady) continue;
    const auto now = Clock::now();
    const int *st = s.h_status.as<int>();
    const double *out = s.h_out.as<double>();
    bool ok = q == hipSuccess;
    if (ok && precision == RX_FP32 && s.h_ctr.as<int>()[CTR_FIX_OVF] != 0) {
      // the FP64 re-check queue overflowed: redo the batch in FP64 (rare;
      // synchronous, outside the queue lock)
      lk.unlock();
      {
        std::lock_guard<std::recursive_mutex> g(G().mu);
        std::string err;
        ok = batch_run(s.batch, RX_FP64, err) == RX_OK;
        ok = ok && hipMemcpyAsync(s.h_out.p, s.batch.d_out.p, sizeof(double) * s.batch.tot_flux,
                                  hipMemcpyDeviceToHost, s.batch.stream) == hipSuccess;
        ok = ok && hipMemcpyAsync(s.h_status.p, s.batch.d_status.p, sizeof(int) * s.batch.n_req,
                                  hipMemcpyDeviceToHost, s.batch.stream) == hipSuccess;
        ok = ok && hipStreamSynchronize(s.batch.stream) == hipSuccess;
      }
      lk.lock();
    }
    for (size_t i = 0; i < s.reqs.size(); i++) {
      const ReqParams &p = s.batch.h_req[i];
      Result &r = results[s.reqs[i].ticket];
      r.done = true;
      const int code = ok ? public_status(st[i]) : RX_ERR_DEVICE;
      r.status = code;
      if (code == RX_OK) r.flux.assign(out + p.flux_off, out + p.flux_off + p.n_in);
      else r.flux.assign(p.n_in, 0.0);
      r.latency_ms = std::chrono::duration<double, std::milli>(now - s.reqs[i].t_submit).count();
      completed++;
    }
    inflight_bytes -= s.bytes;
    inflight_reqs -= (int)s.reqs.size();
    s.reqs.clear();
    s.busy = false;
    any = true;
  }
  if (any) cv.notify_all();
  return any;
}

void rx_server::loop() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  while (!stop.load()) {
    bool progressed = reap(lk);
    launch_ready(lk);
    bool busy = false;
    for (auto &s : slots) busy |= s->busy;
    if (!progressed) {
      if (busy) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
      } else {
        cv.wait_for(lk, std::chrono::milliseconds(5), [&] { return stop.load() || !pending.empty(); });
      }
    }
  }
  // drain whatever is still on the device
  for (auto &s : slots)
    if (s->busy) (void)hipStreamSynchronize(s->batch.stream);
  reap(lk);
}

extern "C" {

rx_server *rx_server_start(int max_inflight, int64_t hbm_budget_bytes, int n_streams, int precision) {
  if (max_inflight <= 0 || n_streams <= 0 || n_streams > 16 || hbm_budget_bytes <= 0) {
    set_last(RX_ERR_ARG, "invalid server configuration");
    return nullptr;
  }
  {
    std::lock_guard<std::recursive_mutex> lk(G().mu);
    std::string err;
    int st = ensure_device(err);
    if (st != RX_OK) {
      set_last(st, err);
      return nullptr;
    }
  }
  auto *s = new rx_server;
  s->max_inflight = max_inflight;
  s->budget = hbm_budget_bytes;
  s->n_streams = n_streams;
  s->precision = precision;
  s->device = G().device;
  for (int i = 0; i < n_streams; i++) s->slots.emplace_back(new Slot);
  s->th = std::thread([s] { s->loop(); });
  return s;
}

int64_t rx_server_submit(rx_server *s, const rx_request *req) {
  if (!s || !req || !req->energy || !req->parameter || req->n_flux <= 0) return -RX_ERR_ARG;
  const int np = n_model_params(req->model);
  if (np < 0) return -RX_ERR_ARG;
  Pending p;
  p.model = req->model;
  p.refl_set = req->refl_set;
  p.energy.assign(req->energy, req->energy + req->n_flux + 1);
  p.params.assign(req->parameter, req->parameter + np);
  if (req->flux_in) p.flux_in.assign(req->flux_in, req->flux_in + req->n_flux);
  p.t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  p.bytes = request_bytes(p);
  p.ticket = s->next_ticket++;
  s->results[p.ticket] = Result();
  s->submitted++;
  const int64_t t = p.ticket;
  s->pending.push_back(std::move(p));
  s->cv.notify_all();
  return t;
}

int rx_server_replay(rx_server *s, const rx_request *req, int n, const double *arrival_s, int64_t *tickets) {
  if (!s || !req || !arrival_s || !tickets || n < 0) return RX_ERR_ARG;
  for (int i = 1; i < n; i++)
    if (!(arrival_s[i] >= arrival_s[i - 1])) return RX_ERR_ARG;
  const auto t0 = Clock::now();
  for (int i = 0; i < n; i++) {
    const auto due = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[i]));
    for (;;) {
      const auto now = Clock::now();
      if (now >= due) break;
      if (due - now > std::chrono::microseconds(200)) std::this_thread::sleep_for(std::chrono::microseconds(100));
      else std::this_thread::yield();
    }
    tickets[i] = rx_server_submit(s, &req[i]);
  }
  return RX_OK;
}

int rx_server_poll(rx_server *s, int64_t ticket, double *flux, int *status, double *latency_ms) {
  if (!s) return -RX_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->results.find(ticket);
  if (it == s->results.end()) return -1;
  if (!it->second.done) return 0;
  if (flux) std::memcpy(flux, it->second.flux.data(), sizeof(double) * it->second.flux.size());
  if (status) *status = it->second.status;
  if (latency_ms) *latency_ms = it->second.latency_ms;
  s->results.erase(it);
  return 1;
}

int rx_server_drain(rx_server *s) {
  if (!s) return RX_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv.wait(lk, [&] { return s->completed >= s->submitted; });
  return RX_OK;
}

int rx_server_stats(rx_server *s, int64_t *out3) {
  if (!s || !out3) return RX_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  out3[0] = s->completed;
  out3[1] = s->batches;
  out3[2] = s->max_concurrent;
  return RX_OK;
}

void rx_server_stop(rx_server *s) {
  if (!s) return;
  s->stop.store(true);
  s->cv.notify_all();
  if (s->th.joinable()) s->th.join();
  delete s;
}

}  // extern "C"
