"""Launch graphs: a batch run repeated with the same kernel arguments is
captured into a hipGraph and replayed.  Every replay (with eager work of other
batches and host synchronisations between the replays) equals the direct
launches; the XSPEC entry point's repeated calls on one grid replay too."""
import numpy as np
import pytest

import relxill_amd as rx
from tests.helpers import (RELLINE_CFG0, abs_err, draw_lp_params, draw_relconv_params, log_grid,
                           reflection_spectrum, rel_err, tables)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tabs():
    rel, lp, T = tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    yield T
    rx.lib.rx_set_graphs(1)


def _requests(n, seed):
    rng = np.random.default_rng(seed)
    lp = draw_lp_params(n, seed=seed + 1)
    cv = draw_relconv_params(n, seed=seed + 2)
    reqs = []
    for k in range(n):
        nb = int(np.exp(rng.uniform(np.log(300), np.log(4000))))
        if k % 3 == 0:
            p = list(RELLINE_CFG0)
            p[4], p[5] = cv[k][3], cv[k][4]
            reqs.append(dict(model=rx.MODEL_RELLINE, energy=log_grid(nb, 0.1, 10.0), params=p))
        elif k % 3 == 1:
            e = log_grid(nb, 0.1, 100.0)
            em = 0.5 * (e[1:] + e[:-1])
            reqs.append(dict(model=rx.MODEL_RELCONV, energy=e, params=cv[k], flux_in=em ** -2.0 * np.diff(e)))
        else:
            reqs.append(dict(model=rx.MODEL_RELLINELP, energy=log_grid(nb, 0.1, 10.0), params=lp[k]))
    return reqs


def _close(y, ref, model):
    if model == rx.MODEL_RELCONV:
        return abs_err(y, ref) < 1e-13
    return rel_err(y, ref) < 1e-12


@pytest.mark.parametrize("prec", [0, 1])
def test_graph_replays_equal_direct_launches(tabs, prec):
    reqs = _requests(30, 71)
    other = rx.Batch(_requests(9, 93))
    b = rx.Batch(reqs)
    try:
        rx.lib.rx_set_graphs(0)
        b.run(prec)
        b.sync()
        ref = [b.output(i)[0] for i in range(len(reqs))]
        assert b.graph_replays() == 0
        rx.lib.rx_set_graphs(1)
        for rep in range(6):
            b.run(prec)
            b.sync()
            out = [b.output(i) for i in range(len(reqs))]
            for i, r in enumerate(reqs):
                assert out[i][1] == 0
                assert _close(out[i][0], ref[i], r["model"]), (rep, i, rel_err(out[i][0], ref[i]))
            # eager work of another batch between the replays
            other.run(1 - prec)
            other.sync()
        # the first graph run captures (the previous run had the same arguments)
        print(f"prec {prec}: {b.graph_replays()} of 6 runs replayed from a graph", flush=True)
        assert b.graph_replays() >= 5
    finally:
        b.close()
        other.close()


def test_profiled_graph_reports_kernel_times(tabs):
    """With profiling on, the graph holds the timing events too."""
    b = rx.Batch(_requests(12, 5))
    rx.lib.rx_profile_enable(1)
    try:
        for _ in range(4):
            b.run(rx.FP32)
            b.sync()
        ms = b.kernel_ms()
        print(f"profiled graph runs: {b.graph_replays()} replays, kernel ms {np.round(ms, 4)}", flush=True)
        assert b.graph_replays() >= 2
        assert ms[1] > 0 and ms[0] > 0
    finally:
        rx.lib.rx_profile_enable(0)
        b.close()


def test_xspec_entry_point_repeats_on_one_grid(tabs):
    e = log_grid(3000, 0.1, 100.0)
    em = 0.5 * (e[1:] + e[:-1])
    f_in = em ** -2.0 * np.diff(e)
    p = [3.0, 3.0, 15.0, 0.9, 30.0, -1.0, 400.0, 0.0]
    rx.lib.rx_set_graphs(0)
    ref = rx.lmodrelconv(e, p, f_in)
    rx.lib.rx_set_graphs(1)
    for k in range(3):
        # three calls on one grid (the second captures, the third replays),
        # then other arguments, and the same again
        for _ in range(3):
            y = rx.lmodrelconv(e, p, f_in)
            assert abs_err(y, ref) < 1e-13, k
        rx.lmodrelline(log_grid(500, 0.1, 10.0), RELLINE_CFG0)
