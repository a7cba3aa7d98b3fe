#!/usr/bin/env python3
"""Benchmark of the relativistic smearing path on one MI355X.

Headline workload (the value): an uneven batch of 4096 relconv requests, each
with its own log energy grid of 300-8000 bins (log-uniform) over 0.1-100 keV,
parameters from the lamp-post priors (a U[-0.998,0.998], incl U[5,85] deg,
Rin U[ISCO,50]) with a broken power law q=3, Rout=400, on the 4096-bin
convolution grid; 25x22x100x20 relline table.  One step = one pass of the whole
device path (setup, relline integration, renormalisation, FFT convolution,
rebin) over the batch, FP32 integrand with FP64 sums over the radii (which
replace the Kahan-compensated FP32 sums: no long FP32 running sum exists),
inputs resident in HBM.

Also reported: the plain FP64 path (one request per launch), the batched
lamp-post relline (1024 requests, 3000 bins), the relxill-style convolution of
a resident reflection set, relxill and relxilllp (10 zones) with the xillver
table interpolated on the device (1024 requests each), the serving loop's p50/p99
latency under Poisson arrivals (relline/relconv/lamp-post mix 40/40/20; the
server's own pacing loop replays a pre-generated stream, at 4k, 20k and 100k
QPS and far above saturation), the
maximum relative error of the FP32 path against the FP64 oracle, the dominant
kernel's roofline figures and the FP64 CPU oracle timed on the host's cores.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
VALU_PEAK_FP32_TFLOPS = 157.3  # MI355X published vector FP32
VALU_PEAK_FP64_TFLOPS = 78.6   # MI355X published vector FP64
KERNELS = ("k_setup", "k_integrate", "k_fixup", "k_finalize", "k_conv", "k_xill", "k_primary")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kerr_rms(a):
    s = -1.0 if a < 0 else 1.0
    z1 = 1.0 + (1.0 - a * a) ** (1 / 3) * ((1.0 + a) ** (1 / 3) + (1.0 - a) ** (1 / 3))
    z2 = np.sqrt(3.0 * a * a + z1 * z1)
    return 3.0 + z2 - s * np.sqrt((3.0 - z1) * (3.0 + z1 + 2 * z2))


def log_grid(n_bins, emin, emax):
    i = np.arange(n_bins + 1, dtype=np.float64)
    return np.exp(i / n_bins * (np.log(emax) - np.log(emin)) + np.log(emin))


def refl_input(e):
    """cutoff power law (Gamma=2, Ecut=300 keV) + Gaussians at 6.4 / 6.97 keV, photons/bin"""
    em = 0.5 * (e[1:] + e[:-1])
    de = np.diff(e)
    lines = 0.02 * np.exp(-0.5 * ((em - 6.4) / 0.05) ** 2) + 0.008 * np.exp(-0.5 * ((em - 6.97) / 0.05) ** 2)
    return (em ** -2.0 * np.exp(-em / 300.0) + lines) * de


def make_cfg3(n_req, seed=5):
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n_req):
        n = int(np.exp(rng.uniform(np.log(300), np.log(8000))))
        a = rng.uniform(-0.998, 0.998)
        incl = rng.uniform(5.0, 85.0)
        rin = rng.uniform(kerr_rms(a), 50.0)
        e = log_grid(n, 0.1, 100.0)
        reqs.append(dict(model=11, energy=e, params=[3.0, 3.0, 15.0, a, incl, rin, 400.0, 0.0],
                         flux_in=refl_input(e)))
    return reqs


def refl_angles(n_incl=10, seed=2):
    """per-angle reflection spectra on 2999 log bins over 0.07-1000 keV: the cfg3
    spectrum with a smooth angle modulation (seeded)"""
    rng = np.random.default_rng(seed)
    e = log_grid(2999, 0.07, 1000.0)
    em = 0.5 * (e[1:] + e[:-1])
    de = np.diff(e)
    base = em ** -2.0 * np.exp(-em / 300.0)
    lines = 0.02 * np.exp(-0.5 * ((em - 6.4) / 0.05) ** 2) + 0.008 * np.exp(-0.5 * ((em - 6.97) / 0.05) ** 2)
    out = np.empty((n_incl, e.size - 1))
    for k in range(n_incl):
        mu = 1.0 - (k + 0.5) / n_incl
        amp = 1.0 + 0.3 * (mu - 0.5) + 0.05 * rng.standard_normal()
        out[k] = (base * em ** (0.05 * (mu - 0.5)) + lines * (1.0 + 0.5 * mu)) * amp * de
    return e, out


def relxill_style(rx, steps):
    """relxill-style angle-resolved relconv: profile on 4096 log bins over
    0.01-1000 keV with 10 angle bins, resident reflection set, output on 3000
    bins over 0.1-100 keV.  Plain FP64 (one request per launch) and a batch of
    1024 parameter sets (FP32 and FP64)."""
    rx.set_conv_grid(4096, 0.01, 1000.0)
    e_ref, refl = refl_angles()
    sid = rx.reflset_create(e_ref, refl)
    e_out = log_grid(3000, 0.1, 100.0)
    rng = np.random.default_rng(7)
    reqs = []
    for _ in range(1024):
        a = rng.uniform(-0.998, 0.998)
        incl = rng.uniform(5.0, 85.0)
        rin = rng.uniform(kerr_rms(a), 50.0)
        reqs.append(dict(model=rx.MODEL_RELXILL_CONV, energy=e_out, params=[3.0, 3.0, 15.0, a, incl, rin, 400.0, 0.0], refl_set=sid))
    lat = []
    for r in reqs[:8]:
        pb = rx.Batch([r])
        d, _ = timed_steps(rx, pb, rx.FP64, 3, 1)
        lat.append(d)
        pb.close()
    b = rx.Batch(reqs)
    b.autotune(rx.FP32)
    dt32, _ = timed_steps(rx, b, rx.FP32, max(3, steps // 2), 1)
    o32 = [b.output(i)[0] for i in range(len(reqs))]
    dt64, _ = timed_steps(rx, b, rx.FP64, 2, 1)
    o64 = [b.output(i)[0] for i in range(len(reqs))]
    b.close()
    rx.lib.rx_reflsets_clear()
    rx.set_conv_grid(4096, 0.00035, 2000.0)
    return dict(plain_fp64_ms_per_spectrum=float(np.mean(lat)) * 1e3,
                batched_fp32_spectra_per_s=len(reqs) / dt32, batched_fp64_spectra_per_s=len(reqs) / dt64,
                max_rel_err_fp32_vs_fp64=max(rel_err(o32[i], o64[i]) for i in range(len(reqs))))


def make_relxilllp_reqs(rx, n_req=1024, seed=10):
    """relxilllp requests: h U[3,100], a, incl, Rin over their ranges, the
    xillver parameters drawn as for relxill, refl_frac 1, no returning
    radiation, output on 3000 bins over 0.1-100 keV"""
    rng = np.random.default_rng(seed)
    e_out = log_grid(3000, 0.1, 100.0)
    reqs = []
    for _ in range(n_req):
        a = rng.uniform(-0.998, 0.998)
        reqs.append(dict(model=rx.MODEL_RELXILLLP, energy=e_out,
                         params=[rng.uniform(3.0, 100.0), 0.0, a, rng.uniform(5.0, 85.0),
                                 rng.uniform(kerr_rms(a), 50.0), 400.0, 0.0, rng.uniform(1.0, 3.4),
                                 rng.uniform(0.0, 4.7), rng.uniform(0.5, 10.0),
                                 float(np.exp(rng.uniform(np.log(5.0), np.log(1000.0)))), 1.0, 0.0, 0.0]))
    return reqs


def relxilllp_batch(rx, steps, n_req=1024, seed=10):
    """relxilllp (lamp post, 10 zones) on the same table and grids
    (make_relxilllp_reqs); FP32 and FP64 batches, plain FP64, the
    FP32-vs-FP64 error and the per-kernel times."""
    reqs = make_relxilllp_reqs(rx, n_req, seed)
    e_out = reqs[0]["energy"]
    b = rx.Batch(reqs)
    b.autotune(rx.FP32)
    dt32, kms32 = timed_steps(rx, b, rx.FP32, max(3, steps // 2), 1)
    o32 = [b.output(i)[0] for i in range(n_req)]
    xv = rx.xilltable_view()
    n_e, n_incl = xv["n_ener"], xv["n_val"][-1]
    n_angles = 0
    n_rows = 0
    for i in range(n_req):
        d = b.dist(i, 10 * n_incl).reshape(10, n_incl)
        ect = b.zones(i, 10)[:, 4].astype(int)
        n_angles += int(np.count_nonzero(d))
        # distinct (Ecut node, angle) pairs over the zones: zones whose shifted
        # Ecut falls between the same table nodes need the same rows
        n_rows += len({(e + k, a) for z in range(10) for a in np.flatnonzero(d[z]) for e in (ect[z],)
                       for k in (0, 1)})
    dt64, _ = timed_steps(rx, b, rx.FP64, 2, 1)
    o64 = [b.output(i)[0] for i in range(n_req)]
    b.close()
    lat = []
    for r in reqs[:8]:
        pb = rx.Batch([r])
        d, _ = timed_steps(rx, pb, rx.FP64, 3, 1)
        lat.append(d)
        pb.close()
    # k_xill algorithmic bytes: per request the distinct table rows its zones
    # need (8 Gamma / A_Fe / logXi corners times the distinct (Ecut node, angle)
    # pairs, n_e float bins each: zones sharing rows read them once from HBM,
    # the repeats come from L2); every zone's FP64 spectrum written
    alg = 8 * n_rows * n_e * 4 + n_req * 10 * 8 * n_e
    t_x = kms32[5] * 1e-3
    return dict(n_requests=n_req, zones=10, mean_angles_used_per_zone=n_angles / n_req / 10,
                rows_read_per_zone_angle=16.0, distinct_rows_per_request=8 * n_rows / n_req,
                k_xill_roofline=dict(bound="hbm", algorithmic_bytes=alg, achieved=alg / t_x / 1e9, peak=HBM_PEAK_GBS,
                                     unit="GB/s", frac=alg / t_x / 1e9 / HBM_PEAK_GBS),
                batched_fp32_spectra_per_s=n_req / dt32,
                batched_fp64_spectra_per_s=n_req / dt64, plain_fp64_ms_per_spectrum=float(np.mean(lat)) * 1e3,
                max_rel_err_fp32_vs_fp64=max(rel_err(o32[i], o64[i]) for i in range(n_req)),
                kernel_ms={k: round(v, 4) for k, v in zip(KERNELS, kms32)})


def make_relxill_reqs(rx, n_req=1024, seed=9):
    """relxill requests: a, incl, Rin and the four xillver parameters drawn
    over their ranges, output on 3000 bins over 0.1-100 keV"""
    rng = np.random.default_rng(seed)
    e_out = log_grid(3000, 0.1, 100.0)
    reqs = []
    for _ in range(n_req):
        a = rng.uniform(-0.998, 0.998)
        incl = rng.uniform(5.0, 85.0)
        rin = rng.uniform(kerr_rms(a), 50.0)
        reqs.append(dict(model=rx.MODEL_RELXILL, energy=e_out,
                         params=[3.0, 3.0, 15.0, a, incl, rin, 400.0, 0.0, rng.uniform(1.0, 3.4),
                                 rng.uniform(0.0, 4.7), rng.uniform(0.5, 10.0),
                                 float(np.exp(rng.uniform(np.log(5.0), np.log(1000.0)))), 3.0]))
    return reqs


def relxill_batch(rx, steps, n_req=1024, seed=9, cpu=True):
    """relxill (broken power law, one zone) on the default synthetic xillver
    table (13 Gamma x 5 A_Fe x 8 logXi x 6 Ecut x 10 inclinations x 2999 bins):
    a batch of parameter sets drawn over the lmodel.dat ranges, output on 3000
    bins over 0.1-100 keV.  FP32 and FP64 batches, plain FP64 (one request per
    launch), the k_xill roofline and the oracle on one host core."""
    rx.xilltable_synthetic()
    reqs = make_relxill_reqs(rx, n_req, seed)
    e_out = reqs[0]["energy"]
    b = rx.Batch(reqs)
    b.autotune(rx.FP32)
    dt32, kms32 = timed_steps(rx, b, rx.FP32, max(3, steps // 2), 1)
    o32 = [b.output(i)[0] for i in range(n_req)]
    xv = rx.xilltable_view()
    n_e, n_incl = xv["n_ener"], xv["n_val"][-1]
    n_angles = sum(int(np.count_nonzero(b.dist(i, n_incl))) for i in range(n_req))
    dt64, _ = timed_steps(rx, b, rx.FP64, 2, 1)
    o64 = [b.output(i)[0] for i in range(n_req)]
    b.close()
    # the primary continuum on a caller grid coarser than the convolution
    # grid (200 bins over 0.1-1000 keV): each output bin sums many
    # convolution bins in one lane
    e_coarse = log_grid(200, 0.1, 1000.0)
    bc = rx.Batch([dict(r, energy=e_coarse) for r in reqs])
    _, kms_c = timed_steps(rx, bc, rx.FP32, 3, 1)
    bc.close()
    lat = []
    for r in reqs[:8]:
        pb = rx.Batch([r])
        d, _ = timed_steps(rx, pb, rx.FP64, 3, 1)
        lat.append(d)
        pb.close()
    # k_xill algorithmic bytes: per request and emission angle of nonzero
    # weight, 16 table corners of n_e float bins read; the FP64 result written
    alg = 16 * n_angles * n_e * 4 + n_req * 8 * n_e
    t_x = kms32[5] * 1e-3
    out = dict(n_requests=n_req, batched_fp32_spectra_per_s=n_req / dt32, batched_fp64_spectra_per_s=n_req / dt64,
               plain_fp64_ms_per_spectrum=float(np.mean(lat)) * 1e3,
               max_rel_err_fp32_vs_fp64=max(rel_err(o32[i], o64[i]) for i in range(n_req)),
               kernel_ms={k: round(v, 4) for k, v in zip(KERNELS, kms32)},
               k_primary_ms_200_bin_grid=round(float(kms_c[6]), 4),
               mean_angles_used=n_angles / n_req,
               k_xill_roofline=dict(bound="hbm", algorithmic_bytes=alg, achieved=alg / t_x / 1e9,
                                    peak=HBM_PEAK_GBS, unit="GB/s", frac=alg / t_x / 1e9 / HBM_PEAK_GBS))
    if cpu:
        import oracle
        rel, lp = rx.tables_view()
        T = oracle.Tables(rel, lp)
        xt = oracle.Xill(xv)
        t0 = time.perf_counter()
        n_cpu = 8
        ref = [oracle.relxill(T, xt, r["params"], e_out)[0] for r in reqs[:n_cpu]]
        t1 = time.perf_counter()
        out["cpu_oracle_single_core_spectra_per_s"] = n_cpu / (t1 - t0)
        out["max_rel_err_fp32_vs_oracle"] = max(rel_err(o32[i], ref[i]) for i in range(n_cpu))
    return out


def make_cfg2(n_req, seed=4):
    rng = np.random.default_rng(seed)
    e = log_grid(3000, 0.1, 10.0)
    reqs = []
    for _ in range(n_req):
        a = rng.uniform(-0.998, 0.998)
        h = rng.uniform(3.0, 100.0)
        incl = rng.uniform(5.0, 85.0)
        rin = rng.uniform(kerr_rms(a), 50.0)
        reqs.append(dict(model=2, energy=e, params=[6.4, h, a, incl, rin, 400.0, 0.0, 0.0, 2.0, 0.0]))
    return reqs


def rel_err(y, yref, floor=1e-10):
    m = np.abs(yref).max()
    if m == 0:
        return float(np.abs(y).max())
    sel = np.abs(yref) > floor * m
    return float(np.max(np.abs(y[sel] - yref[sel]) / np.abs(yref[sel])))


def timed_steps(rx, b, prec, steps, warmup):
    for _ in range(warmup):
        b.run(prec)
    b.sync()
    rx.lib.rx_device_synchronize()
    kms = np.zeros(len(KERNELS))
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run(prec)
        b.sync()
        kms += b.kernel_ms()
    rx.lib.rx_device_synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) / steps, kms / steps


def serving_pool(n=2048, seed=6):
    """cfg4 requests: relline / relconv / lamp-post relline at 40/40/20, grids
    of 300-8000 bins (log-uniform), parameters from the lamp-post priors."""
    rng = np.random.default_rng(seed)
    pool = []
    for _ in range(n):
        nb = int(np.exp(rng.uniform(np.log(300), np.log(8000))))
        a = rng.uniform(-0.998, 0.998)
        incl = rng.uniform(5.0, 85.0)
        rin = rng.uniform(kerr_rms(a), 50.0)
        u = rng.uniform()
        if u < 0.4:
            pool.append(dict(model=1, energy=log_grid(nb, 0.1, 10.0),
                             params=[6.4, 3.0, 3.0, 15.0, a, incl, rin, 400.0, 0.0, 0.0]))
        elif u < 0.8:
            e = log_grid(nb, 0.1, 100.0)
            pool.append(dict(model=11, energy=e, params=[3.0, 3.0, 15.0, a, incl, rin, 400.0, 0.0],
                             flux_in=refl_input(e)))
        else:
            pool.append(dict(model=2, energy=log_grid(nb, 0.1, 10.0),
                             params=[6.4, rng.uniform(3.0, 100.0), a, incl, rin, 400.0, 0.0, 0.0, 2.0, 0.0]))
    return pool


def serve(rx, pool, qps, seconds=None, n_total=None, seed=6, warmup_s=0.25):
    """Poisson arrivals at qps (seconds long, or n_total requests) drawn from
    the pre-generated pool, replayed by the server's native pacing loop (so
    the client does not limit the offered load); p50 / p99 of submit ->
    completion and the completed spectra per second of wall time.  The
    measured stream follows a warm-up stream of warmup_s seconds at the same
    rate on the same server (its slots' first buffer allocations), which is
    not counted."""
    rng = np.random.default_rng(seed)
    n = n_total if n_total else int(qps * seconds * 1.2) + 16
    t = np.cumsum(rng.exponential(1.0 / qps, size=n))
    if seconds and not n_total:
        t = t[t <= seconds]
    reqs = [pool[i % len(pool)] for i in range(t.size)]
    srv = rx.Server(max_inflight=8192, hbm_budget=32 << 30, n_streams=4, precision=rx.FP32)
    if warmup_s > 0:
        tw = np.cumsum(rng.exponential(1.0 / qps, size=int(qps * warmup_s * 1.2) + 16))
        tw = tw[tw <= warmup_s]
        srv.replay([pool[(i * 7) % len(pool)] for i in range(tw.size)], tw)
        srv.drain()
    st0 = srv.stats()
    t0 = time.perf_counter()
    tickets = srv.replay(reqs, t)
    srv.drain()
    wall = time.perf_counter() - t0
    lats, bad = np.empty(len(tickets)), 0
    for k, tk in enumerate(tickets):
        r = srv.poll(tk)
        lats[k] = r[2]
        bad += r[1] != 0
    st = srv.stats()
    srv.stop()
    nb = st["batches"] - st0["batches"]
    return dict(qps_offered=qps, submitted=len(tickets), failed=int(bad),
                p50_ms=float(np.percentile(lats, 50)), p99_ms=float(np.percentile(lats, 99)),
                spectra_per_s=len(tickets) / wall, mean_batch=len(tickets) / max(1, nb),
                batches=nb, max_concurrent_batches=st["max_concurrent_batches"], warmup_s=warmup_s)


def host_cpus():
    """What the host offers this process: logical CPUs (nproc), physical cores
    (distinct (physical id, core id) pairs of /proc/cpuinfo), the CPU model,
    the CPUs this process may run on (affinity) and its cgroup CPU quota."""
    info = dict(nproc=os.cpu_count() or 1, model=None, physical_cores=None, affinity=None, cgroup_quota=None)
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        cores, phys = set(), None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and info["model"] is None:
                info["model"] = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                cores.add((phys, v))
        info["physical_cores"] = len(cores) or None
    except OSError:
        pass
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            info["cgroup_quota"] = float(q) / float(per)
    except (OSError, ValueError):
        pass
    return info


def cpu_threads(info):
    """Threads for the all-cores leg: the CPUs this process is granted (the
    smaller of its affinity set and its cgroup quota), else nproc."""
    n = info["affinity"] or info["nproc"]
    if info["cgroup_quota"]:
        n = min(n, max(1, int(info["cgroup_quota"])))
    return max(1, n)


def cpu_baseline(reqs, n_threads, gpu_out):
    """The FP64 oracle on the host: single core on 16 requests, then the sample
    on n_threads threads; also the reference algorithm's work counters over
    the sample (integrand evaluations and the rest), the FLOP count of the
    roofline."""
    import oracle
    import relxill_amd as rx
    rel, lp = rx.tables_view()
    T = oracle.Tables(rel, lp)
    jobs = [(r["model"], r["params"], r["energy"], r["flux_in"]) for r in reqs]
    # single core
    t0 = time.perf_counter()
    outs1, st1 = oracle.eval_batch_mt(T, jobs[:16], 1)
    t1 = time.perf_counter()
    single = 16 / (t1 - t0)
    t0 = time.perf_counter()
    outs, st = oracle.eval_batch_mt(T, jobs, n_threads)
    t1 = time.perf_counter()
    allc = len(jobs) / (t1 - t0)
    err = max(rel_err(gpu_out[i], outs[i]) for i in range(len(jobs)) if st[i] == 0)
    return dict(single_core=single, all_cores=allc, cpu_seconds=(t1 - t0) * n_threads,
                max_rel_err_fp32_vs_oracle=err, n=len(jobs))


def oracle_work(reqs, n_threads):
    """The reference algorithm's work counters summed over every request of
    the batch (oracle restatement, counting only)."""
    import oracle
    import relxill_amd as rx
    rel, lp = rx.tables_view()
    T = oracle.Tables(rel, lp)
    jobs = [(r["model"], r["params"], r["energy"], r.get("flux_in")) for r in reqs]
    _, _, stats = oracle.eval_batch_mt(T, jobs, n_threads, stats=True)
    return stats


def oracle_errors(reqs, outs, n_threads):
    """max relative error per bin of outs against the FP64 oracle on the same
    requests"""
    import oracle
    import relxill_amd as rx
    rel, lp = rx.tables_view()
    T = oracle.Tables(rel, lp)
    jobs = [(r["model"], r["params"], r["energy"], r.get("flux_in")) for r in reqs]
    ref, st = oracle.eval_batch_mt(T, jobs, n_threads)
    assert all(x == 0 for x in st)
    return max(rel_err(outs[i], ref[i]) for i in range(len(reqs)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-requests", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--cpu-sample", type=int, default=2048)
    ap.add_argument("--shape", type=str, default="",
                    help="fixed work shape 'seg_bins,waves_per_cu' instead of autotuning")
    ap.add_argument("--profile-only", action="store_true",
                    help="only the headline batch (FP32), for profiler runs")
    ap.add_argument("--relxill-only", action="store_true",
                    help="only the relxill batch (for profiler runs)")
    ap.add_argument("--profile-batch", choices=["relxill", "relxilllp"], default="",
                    help="one 1024-request relxill or relxilllp batch at the --shape work shape, FP32, "
                         "--warmup untimed and --steps timed runs, nothing else (for profiler runs)")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="run a plain copy kernel of this many MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(requests shard across GPUs with no exchange between them)", file=sys.stderr)
        sys.exit(2)

    import relxill_amd as rx
    rx.tables_synthetic(1, 25, 22, 100, 20)
    rx.lptable_synthetic(3, 20, 100, 100)
    rx.set_conv_grid(4096, 0.00035, 2000.0)
    if rx.lib.rx_init(0) != 0:
        raise SystemExit("no GPU: " + rx.lib.rx_last_error().decode())
    rx.lib.rx_profile_enable(1)
    if args.copy_calibration > 0:
        nbytes = int(args.copy_calibration * (1 << 20))
        ms = np.zeros(1)
        assert rx.lib.rx_copy_test(nbytes, 5, ms.ctypes.data_as(rx.capi._D)) == 0
        print(json.dumps({"copy_bytes": nbytes, "copy_ms": float(ms[0]),
                          "copy_GBps": 2 * nbytes / ms[0] / 1e6}))
        return

    if args.profile_batch:
        rx.xilltable_synthetic()
        make = make_relxill_reqs if args.profile_batch == "relxill" else make_relxilllp_reqs
        b = rx.Batch(make(rx))
        seg, wpc = (int(x) for x in (args.shape or "128,12").split(","))
        b.set_shape(seg, wpc)
        dt, kms = timed_steps(rx, b, rx.FP32, args.steps, args.warmup)
        print(json.dumps({"batch": args.profile_batch, "ms_per_step": dt * 1e3, "kernel_ms": kms.tolist()}))
        b.close()
        return

    if args.relxill_only:
        r = relxill_batch(rx, args.steps, cpu=False)
        r["relxilllp"] = relxilllp_batch(rx, args.steps)
        print(json.dumps(r))
        return

    log("building cfg3 batch ...")
    reqs = make_cfg3(args.n_requests)
    nbins = np.array([r["energy"].size - 1 for r in reqs])
    b = rx.Batch(reqs)
    if args.shape:
        seg, wpc = (int(x) for x in args.shape.split(","))
        b.set_shape(seg, wpc)
    else:
        seg, wpc = b.autotune(rx.FP32)
    log(f"work shape: {seg} bins per item, {wpc} waves per CU")
    dt32, kms32 = timed_steps(rx, b, rx.FP32, args.steps, args.warmup)
    value = len(reqs) / dt32
    log(f"cfg3 fp32: {dt32 * 1e3:.2f} ms/step, {value:.1f} spectra/s, kernels ms {np.round(kms32, 3)}")
    if args.profile_only:
        print(json.dumps({"ms_per_step": dt32 * 1e3, "kernel_ms": kms32.tolist(), "value": value}))
        return
    out32 = [b.output(i)[0] for i in range(len(reqs))]
    work32 = b.counters()
    log(f"cfg3 fp32 work: {work32}")

    dt64b, kms64b = timed_steps(rx, b, rx.FP64, max(2, args.steps // 3), 1)
    out64 = [b.output(i)[0] for i in range(len(reqs))]
    err32_64 = max(rel_err(out32[i], out64[i]) for i in range(len(reqs)))
    log(f"cfg3 fp64 batched: {dt64b * 1e3:.2f} ms/step; max rel err fp32 vs fp64 {err32_64:.3e}")
    b.close()

    # plain FP64: one request per launch (device-resident single-request batches)
    lat = []
    for r in reqs[:16]:
        pb = rx.Batch([r])
        d, _ = timed_steps(rx, pb, rx.FP64, 3, 1)
        lat.append(d)
        pb.close()
    plain64 = 1.0 / float(np.mean(lat))
    log(f"plain fp64: {np.mean(lat) * 1e3:.3f} ms/spectrum -> {plain64:.1f} spectra/s")
    # XSPEC entry point (host arrays in/out, includes PCIe and launch overhead)
    t0 = time.perf_counter()
    for r in reqs[:16]:
        rx.lmodrelconv(r["energy"], r["params"], r["flux_in"])
    lmod_ms = (time.perf_counter() - t0) / 16 * 1e3

    # cfg2: lamp-post relline batch
    b2 = rx.Batch(make_cfg2(1024))
    b2.autotune(rx.FP32)
    dt2, kms2 = timed_steps(rx, b2, rx.FP32, max(3, args.steps // 2), 1)
    o2_32 = [b2.output(i)[0] for i in range(1024)]
    timed_steps(rx, b2, rx.FP64, 1, 0)
    o2_64 = [b2.output(i)[0] for i in range(1024)]
    err2 = max(rel_err(o2_32[i], o2_64[i]) for i in range(1024))
    cfg2 = dict(spectra_per_s=1024 / dt2, ms_per_step=dt2 * 1e3, max_rel_err_fp32_vs_fp64=err2)
    cfg2_reqs = make_cfg2(1024)
    log(f"cfg2 lp relline fp32: {cfg2}")
    b2.close()

    # cfg1: relxill-style angle-resolved convolution
    cfg1 = relxill_style(rx, args.steps)
    log(f"cfg1 relxill-style: {cfg1}")
    # relxill with the xillver table on the device
    relxill = relxill_batch(rx, args.steps, cpu=not args.no_cpu)
    log(f"relxill batch: {relxill}")
    relxilllp = relxilllp_batch(rx, args.steps)
    log(f"relxilllp batch: {relxilllp}")

    serving = serving_load = None
    if not args.no_serve:
        pool = serving_pool()
        serving = serve(rx, pool, 4000.0, seconds=args.serve_seconds)
        log(f"serving at 4k QPS: {serving}")
        serving_load = []
        for q in (20000.0, 100000.0):
            r = serve(rx, pool, q, seconds=min(args.serve_seconds, 2.0), seed=16)
            log(f"serving at {q:.0f} QPS: {r}")
            serving_load.append(r)
        # saturation: far more offered than the server can take
        r = serve(rx, pool, 2.0e6, n_total=60000, seed=26)
        log(f"serving saturated (2M QPS offered): {r}")
        serving_load.append(r)

    # roofline of the dominant kernel (relline integration): algorithmic bytes
    # per spectrum = table corners 68,800 B + grid/in/out 16 N + 8 B
    alg_bytes = float(np.sum(68800 + 16 * nbins + 8))
    t_int = kms32[1] * 1e-3
    achieved = alg_bytes / t_int / 1e9
    # measured HBM bytes per launch of the same kernel and work shape, from the
    # committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated on a copy)
    # the kernel is bound by VALU issue, not HBM: the same summary's VALU-busy
    # fraction (SQ_ACTIVE_INST_VALU over GPU-busy cycles) is reported beside it
    # (the summary is taken at one work shape, named beside the figure: the
    # traffic of the list's other shapes differs by the re-staged node rows)
    traffic, traffic_src, valu_busy, traffic_shape = None, None, None, None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_relconv_batch4096.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        ki = tj.get("kernels", {}).get("k_integrate<1, 0>")
        if ki and len(reqs) == 4096:
            traffic = round(ki["read_bytes"] + ki["write_bytes"])
            traffic_src = tj.get("source")
            traffic_shape = tj.get("shape")
            valu_busy = ki.get("valu_busy")
    roofline = dict(bound="hbm", achieved=round(achieved, 3), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=achieved / HBM_PEAK_GBS, traffic=traffic, traffic_source=traffic_src,
                    traffic_shape=traffic_shape, valu_busy=valu_busy,
                    algorithmic_bytes=round(alg_bytes),
                    kernel="k_integrate<1, 0>", kernel_ms=round(kms32[1], 4),
                    kernel_ms_all={k: round(v, 4) for k, v in zip(KERNELS, kms32)})

    cpu = None
    host = host_cpus()
    max_err_oracle = None
    if not args.no_cpu:
        ncpu = cpu_threads(host)
        sample = list(range(min(args.cpu_sample, len(reqs))))
        cb = cpu_baseline([reqs[i] for i in sample], ncpu, [out32[i] for i in sample])
        log(f"cpu baseline: {cb} host {host}")
        ncores = host["physical_cores"] or host["nproc"]
        cpu = dict(value=round(cb["all_cores"], 2), unit="spectra/s", cores=ncpu, kind="port",
                   sample=f"first {cb['n']} requests of the same batch, FP64 oracle on {ncpu} threads "
                          f"(the CPUs this process is granted), {cb['cpu_seconds']:.1f} s of CPU work; "
                          f"single core {cb['single_core']:.2f} spectra/s",
                   single_core=round(cb["single_core"], 3),
                   host=dict(cpu_model=host["model"], nproc=host["nproc"], physical_cores=host["physical_cores"],
                             affinity=host["affinity"], cgroup_quota=host["cgroup_quota"]),
                   all_host_cores_extrapolated=round(cb["single_core"] * ncores, 1))
        max_err_oracle = cb["max_rel_err_fp32_vs_oracle"]
        # cfg2: a sample of the lamp-post batch against the oracle
        n2 = 64
        cfg2["max_rel_err_fp32_vs_oracle"] = oracle_errors(cfg2_reqs[:n2], o2_32[:n2], ncpu)
        cfg2["oracle_sample"] = n2
        log(f"cfg2 fp32 vs oracle on {n2} requests: {cfg2['max_rel_err_fp32_vs_oracle']:.3e}")
        # algorithmic FLOPs of the integration: the reference algorithm's
        # integrand evaluations over the whole batch (20 FLOP each: the g*
        # map, the node lerp, E^3, the 1/sqrt(g*(1-g*)) weight and the
        # products, with sqrt and division counted as one FLOP)
        work = oracle_work(reqs, ncpu)
        flops = 20.0 * work["relb_func_calls"]
        t_int = kms32[1] * 1e-3
        roofline_valu = dict(algorithmic_flops=flops, integrand_evaluations=work["relb_func_calls"],
                             flop_per_evaluation=20, achieved=round(flops / t_int / 1e12, 4),
                             peak=VALU_PEAK_FP32_TFLOPS, unit="TFLOP/s",
                             frac=flops / t_int / 1e12 / VALU_PEAK_FP32_TFLOPS,
                             frac_of_fp64_peak=flops / t_int / 1e12 / VALU_PEAK_FP64_TFLOPS)
        roofline = dict(bound="valu", achieved=roofline_valu["achieved"], peak=VALU_PEAK_FP32_TFLOPS,
                        unit="TFLOP/s", frac=roofline_valu["frac"], traffic=roofline["traffic"],
                        kernel=roofline["kernel"], kernel_ms=roofline["kernel_ms"],
                        valu=roofline_valu, hbm={k: roofline[k] for k in ("achieved", "peak", "unit", "frac",
                                                                          "traffic", "traffic_source",
                                                                          "traffic_shape", "algorithmic_bytes")},
                        valu_busy=roofline["valu_busy"], kernel_ms_all=roofline["kernel_ms_all"])
        log(f"roofline: {roofline}")

    res = {
        "metric": "spectra/sec batched FP32 vs plain FP64 on 1 GPU @<=1e-4 rel err; p99 latency",
        "value": round(value, 2),
        "unit": "spectra/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt32 * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "uneven relconv batch: 4096 requests, own log grids of 300-8000 bins "
                        "(log-uniform) over 0.1-100 keV, 4096-bin convolution grid 0.00035-2000 keV, "
                        "25x22x100x20 relline table, FP32 integrand with FP64 sums over radii (in place of Kahan)",
            "n_requests": len(reqs), "mean_bins": float(nbins.mean()), "seg_bins": seg,
            "waves_per_cu": wpc,
        },
        "plain_fp64_spectra_per_s": round(plain64, 2),
        "ratio_vs_plain_fp64": round(value / plain64, 2),
        "batched_fp64_spectra_per_s": round(len(reqs) / dt64b, 2),
        "lmodrelconv_ms_per_call": round(lmod_ms, 3),
        "max_rel_err_fp32_vs_fp64_gpu": err32_64,
        "fp32_work": work32,
        "max_rel_err_fp32_vs_oracle": max_err_oracle,
        "lp_relline_batch": cfg2,
        "relxill_style_conv": cfg1,
        "relxill_batch": relxill,
        "relxilllp_batch": relxilllp,
        "serving": serving,
        "serving_load_curve": serving_load,
        "serving_saturation_spectra_per_s": serving_load[-1]["spectra_per_s"] if serving_load else None,
        "p99_latency_ms": serving["p99_ms"] if serving else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        res["ratio_vs_cpu_all_cores"] = round(value / cpu["value"], 1)
        res["ratio_vs_cpu_all_host_cores_extrapolated"] = round(value / cpu["all_host_cores_extrapolated"], 1)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
