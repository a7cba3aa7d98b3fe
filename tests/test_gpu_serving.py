"""The serving workload on the GPU: Poisson arrivals at a fixed rate of a
relline / relconv / lamp-post relline mix at 40/40/20 with grids of 300-8000
bins (log-uniform), tables and inputs resident, FP32; replayed by the
server's native pacing loop for about two seconds.  Every served result
equals its solo run (only the FP64 order of the per-chunk sums may differ),
and 16 requests of each model match the FP64 oracle within 1e-4 per bin."""
import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import abs_err, log_grid, rel_err, tables

pytestmark = pytest.mark.gpu

QPS = 2500.0
SECONDS = 2.0


@pytest.fixture(scope="module")
def tabs():
    rel, lp, T = tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    return T


def _stream(seed=6):
    """Arrival times and requests of the serving workload."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / QPS, size=int(QPS * SECONDS * 1.3)))
    t = t[t <= SECONDS]
    reqs = []
    for _ in range(t.size):
        nb = int(np.exp(rng.uniform(np.log(300), np.log(8000))))
        a = rng.uniform(-0.998, 0.998)
        incl = rng.uniform(5.0, 85.0)
        rin = rng.uniform(oracle.lib.orc_kerr_rms(a), 50.0)
        u = rng.uniform()
        if u < 0.4:
            reqs.append(dict(model=rx.MODEL_RELLINE, energy=log_grid(nb, 0.1, 10.0),
                             params=[6.4, 3.0, 3.0, 15.0, a, incl, rin, 400.0, 0.0, 0.0]))
        elif u < 0.8:
            e = log_grid(nb, 0.1, 100.0)
            em = 0.5 * (e[1:] + e[:-1])
            reqs.append(dict(model=rx.MODEL_RELCONV, energy=e, params=[3.0, 3.0, 15.0, a, incl, rin, 400.0, 0.0],
                             flux_in=(em ** -2.0 * np.exp(-em / 300.0) + 0.02 * np.exp(-0.5 * ((em - 6.4) / 0.05) ** 2))
                             * np.diff(e)))
        else:
            reqs.append(dict(model=rx.MODEL_RELLINELP, energy=log_grid(nb, 0.1, 10.0),
                             params=[6.4, rng.uniform(3.0, 100.0), a, incl, rin, 400.0, 0.0, 0.0, 2.0, 0.0]))
    return t, reqs


def test_serving_workload_against_solo_runs_and_oracle(tabs):
    t, reqs = _stream()
    models = [r["model"] for r in reqs]
    assert len(reqs) > 4000
    for m, share in ((rx.MODEL_RELLINE, 0.4), (rx.MODEL_RELCONV, 0.4), (rx.MODEL_RELLINELP, 0.2)):
        assert abs(models.count(m) / len(reqs) - share) < 0.05
    srv = rx.Server(max_inflight=4096, hbm_budget=4 << 30, n_streams=4, precision=rx.FP32)
    try:
        tickets = srv.replay(reqs, t)
        srv.drain()
        got = [srv.poll(tk) for tk in tickets]
        st = srv.stats()
    finally:
        srv.stop()
    lat = np.array([g[2] for g in got])
    print(f"served {len(reqs)} requests in {st['batches']} batches (max {st['max_concurrent_batches']} at once); "
          f"latency p50 {np.percentile(lat, 50):.3f} ms p99 {np.percentile(lat, 99):.3f} ms", flush=True)
    assert st["completed"] == len(reqs)
    assert all(g[1] == 0 for g in got)
    # every result equals its solo run
    worst_line, worst_conv = 0.0, 0.0
    for r, (y, _, _) in zip(reqs, got):
        solo, sst = rx.eval_model(r["model"], r["params"], r["energy"], flux_in=r.get("flux_in"), precision=rx.FP32)
        assert sst == 0
        if r["model"] == rx.MODEL_RELCONV:
            worst_conv = max(worst_conv, abs_err(y, solo))
        else:
            worst_line = max(worst_line, rel_err(y, solo))
    print(f"served vs solo: line models {worst_line:.3e} per bin, relconv {worst_conv:.3e} of the maximum", flush=True)
    assert worst_line < 1e-12 and worst_conv < 1e-13
    # 16 requests of each model against the FP64 oracle
    _, _, T = tables()
    idx = []
    for m in (rx.MODEL_RELLINE, rx.MODEL_RELCONV, rx.MODEL_RELLINELP):
        idx += [i for i, r in enumerate(reqs) if r["model"] == m][:16]
    jobs = [(reqs[i]["model"], reqs[i]["params"], reqs[i]["energy"], reqs[i].get("flux_in")) for i in idx]
    ref, rst = oracle.eval_batch_mt(T, jobs, 8)
    assert all(s == 0 for s in rst)
    errs = [rel_err(got[i][0], ref[k]) for k, i in enumerate(idx)]
    print(f"served vs FP64 oracle (48 requests): max per-bin error {max(errs):.3e}", flush=True)
    assert max(errs) < 1e-4
