// nthcomp.hip -- the nthcomp primary continuum on the GPU.
//
// spec_nthcomp (src/Xillspec.cpp:243-253): thermal Comptonization of a
// disc-blackbody seed (kT_bb = 0.05 keV, get_nthcomp_param,
// src/relutility.c:625-632) as computed by c_donthcomp (src/donthcomp.c:
// 683-793), with the solve and the bin integrals of rx_nthcomp.h.  Two
// kernels:
//   k_nthcomp        one 256-thread workgroup per (Gamma, kTe): the Kompaneets
//                    equation on the log photon grid x_j = 1e-4 T_bb 10^(0.02 j)
//                    (<= 899 points) -- grid, Cooper coefficients, escape
//                    probabilities, the disc seed (five-point Gauss per seed
//                    bin) and the tridiagonal coefficients in parallel, the
//                    tridiagonal solve (forward elimination and back
//                    substitution) by one thread -- written to HBM as
//                    (x, E F_E, jmax);
//   k_nthcomp_bins   one workgroup per spectrum: E F_E interpolated at every
//                    edge of the caller grid (the edge's bracket by binary
//                    search, which finds the bracket of the reference's
//                    monotone walk), divided by E^2, trapezoid per bin, times
//                    the normalisation E F_E(1 keV, source) = 1.
// Every operation keeps the reference's order (contraction off), so results
// differ from the FP64 oracle only by the last-ulp differences of the device
// pow / exp / log.
#include <hip/hip_runtime.h>

#include "engine.h"
#include "rx_math.h"
#include "rx_nthcomp.h"

namespace rx {

namespace {

constexpr int NTH_THREADS = 256;

// One Kompaneets solve per workgroup, written to HBM as (x, E F_E, jmax).
__global__ __launch_bounds__(NTH_THREADS) void k_nthcomp(const NthJob *__restrict__ jobs, double *__restrict__ sol_x,
                                                         double *__restrict__ sol_s, int *__restrict__ sol_n) {
  __shared__ NthLds L;
  const int tid = threadIdx.x;
  const NthJob jb = jobs[blockIdx.x];
  const int jmax = nth_solve(L, jb.kte, jb.gam, jb.ktbb, tid, NTH_THREADS);
  double *xo = sol_x + (size_t)blockIdx.x * (NTH_N + 1);
  double *so = sol_s + (size_t)blockIdx.x * NTH_N;
  for (int j = tid; j <= jmax; j += NTH_THREADS) xo[j] = L.x[j];
  for (int j = tid; j < NTH_N; j += NTH_THREADS) so[j] = L.c[j];
  if (tid == 0) sol_n[blockIdx.x] = jmax;
}

__global__ __launch_bounds__(NTH_THREADS) void k_nthcomp_bins(const double *__restrict__ ener, int n,
                                                              const double *__restrict__ shift,
                                                              const double *__restrict__ sol_x,
                                                              const double *__restrict__ sol_s,
                                                              const int *__restrict__ sol_n, double *__restrict__ out) {
#pragma clang fp contract(off)
  const int s = blockIdx.x;
  const double *x = sol_x + (size_t)s * (NTH_N + 1);
  const double *spt = sol_s + (size_t)s * NTH_N;
  const int nth = sol_n[s];
  const double z = 1 / shift[s] - 1;
  const double normfac = nthcomp_normfac(x, spt, nth, z);
  for (int i = threadIdx.x; i < n; i += NTH_THREADS) {
    const double p0 = nthcomp_edge(x, spt, nth, ener[i], z);
    const double p1 = nthcomp_edge(x, spt, nth, ener[i + 1], z);
    out[(size_t)s * n + i] = nthcomp_bin(ener[i], ener[i + 1], p0, p1, normfac);
  }
}

}  // namespace

int launch_nthcomp(const NthJob *jobs, int n_jobs, double *sol_x, double *sol_s, int *sol_n, void *stream) {
  if (n_jobs <= 0) return 0;
  hipLaunchKernelGGL(k_nthcomp, dim3(n_jobs), dim3(NTH_THREADS), 0, (hipStream_t)stream, jobs, sol_x, sol_s,
                     sol_n);
  return (int)hipGetLastError();
}

// rx_nthcomp: n_spec spectra on one caller grid (host arrays in and out).
int nthcomp_spectra(const double *ener, int n_flux, int n_spec, const double *kte, const double *gam,
                    const double *ener_shift, double *flux, std::string &err) {
  for (int s = 0; s < n_spec; s++) {
    // the solve's grid needs kTe well above the seed and a finite Thomson depth
    if (!(kte[s] >= 0.1 && kte[s] <= 1e4) || !(gam[s] > 1.0 && gam[s] < 10.0) ||
        !(ener_shift[s] > 0.0 && ener_shift[s] < 1e3)) {
      err = "nthcomp parameters out of range (kTe 0.1-1e4 keV, 1 < Gamma < 10, energy shift > 0)";
      return RX_ERR_ARG;
    }
  }
  for (int k = 0; k < n_flux; k++)
    if (!(ener[k + 1] > ener[k]) || !(ener[k] > 0.0)) {
      err = "energy grid must be positive and ascending";
      return RX_ERR_ARG;
    }
  std::vector<NthJob> jobs(n_spec);
  for (int s = 0; s < n_spec; s++) jobs[s] = NthJob{kte[s], gam[s], 0.05};
  DevBuf d_jobs, d_e, d_shift, d_x, d_s, d_n, d_out;
  if (d_jobs.ensure(sizeof(NthJob) * n_spec) || d_e.ensure(sizeof(double) * (n_flux + 1)) ||
      d_shift.ensure(sizeof(double) * n_spec) || d_x.ensure(sizeof(double) * (NTH_N + 1) * n_spec) ||
      d_s.ensure(sizeof(double) * NTH_N * n_spec) || d_n.ensure(sizeof(int) * n_spec) ||
      d_out.ensure(sizeof(double) * (size_t)n_flux * n_spec)) {
    err = "device allocation failed";
    return RX_ERR_NOMEM;
  }
  int e = (int)hipMemcpy(d_jobs.p, jobs.data(), sizeof(NthJob) * n_spec, hipMemcpyHostToDevice);
  e |= (int)hipMemcpy(d_e.p, ener, sizeof(double) * (n_flux + 1), hipMemcpyHostToDevice);
  e |= (int)hipMemcpy(d_shift.p, ener_shift, sizeof(double) * n_spec, hipMemcpyHostToDevice);
  e |= launch_nthcomp(d_jobs.as<NthJob>(), n_spec, d_x.as<double>(), d_s.as<double>(), d_n.as<int>(), nullptr);
  hipLaunchKernelGGL(k_nthcomp_bins, dim3(n_spec), dim3(NTH_THREADS), 0, nullptr, d_e.as<double>(), n_flux,
                     d_shift.as<double>(), d_x.as<double>(), d_s.as<double>(), d_n.as<int>(), d_out.as<double>());
  e |= (int)hipGetLastError();
  e |= (int)hipMemcpy(flux, d_out.p, sizeof(double) * (size_t)n_flux * n_spec, hipMemcpyDeviceToHost);
  if (e) {
    (void)hipGetLastError();
    err = "device error in the nthcomp kernels";
    return RX_ERR_DEVICE;
  }
  return RX_OK;
}

}  // namespace rx
