"""The six-parameter xillver table on the GPU (the outer lerp of k_xill1,
k_xillz and k_rrad_corr, interp_6d_tab_incl / interp_6d_tab,
src/xilltable.c:997-1045) against the oracle: relxillCp, relxilllpCp (its
zones and returning-radiation corrections) and xillverCp on a synthetic
table Gamma x A_Fe x logXi x kTe x Dens x Incl, FP64 and FP32, and the
XSPEC entry points.  Spectra are compared with conv_err (relative to the
bin, or to 1e-5 of the maximum below that: the FP64 round-off of the
convolutions)."""
import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import conv_err, log_grid, rel_err, tables
from tests.test_gpu_relxill_cp import draw_cp, draw_lpcp

pytestmark = pytest.mark.gpu

TOL64 = 1e-8
TOL32 = 1e-4


@pytest.fixture(scope="module")
def cp6():
    _, _, T = tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    rx.xillcptable_synthetic_dens(11, 4, 5, 3, 5, 5, 10, 2999)
    x6 = rx.xilltable_view(nthcomp=True)
    assert x6["n_par"] == 6
    yield T, oracle.Xill(x6)
    rx.xillcptable_synthetic()


def _run(b, n, prec):
    b.run(prec)
    b.sync()
    return [b.output(i) for i in range(n)]


@pytest.mark.parametrize("model", [rx.MODEL_RELXILLCP, rx.MODEL_RELXILLLPCP])
def test_six_parameter_sweep(cp6, model):
    """8 parameter sets over the lmodel.dat ranges (logN 15-19 interpolated on
    the table's density axis, Gamma on its outer axis; relxilllpCp with and
    without returning radiation), one batch, FP64 and FP32."""
    T, x6 = cp6
    om = oracle.MOD_RELXILLCP if model == rx.MODEL_RELXILLCP else oracle.MOD_RELXILLLPCP
    rng = np.random.default_rng(61 if model == rx.MODEL_RELXILLCP else 62)
    e = log_grid(1200, 0.1, 100.0)
    ps = []
    for k in range(8):
        p = draw_cp(rng) if model == rx.MODEL_RELXILLCP else draw_lpcp(rng, returnrad=float(k % 2))
        p[10 if model == rx.MODEL_RELXILLCP else 8] = rng.uniform(15.0, 19.0)
        ps.append(p)
    refs = [oracle.relxill_model(om, T, x6, p, e) for p in ps]
    assert sum(r[2] == 0 for r in refs) >= 5
    b = rx.Batch([dict(model=model, energy=e, params=p) for p in ps])
    try:
        for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
            outs = _run(b, len(ps), prec)
            worst = 0.0
            for i, (ref, _, st) in enumerate(refs):
                y, st_g = outs[i]
                assert (st_g == 0) == (st == 0), (i, ps[i], st, st_g)
                if st == 0:
                    worst = max(worst, conv_err(y, ref))
            print(f"model {model} six-parameter table prec {prec}: worst {worst:.3e}", flush=True)
            assert worst < tol
    finally:
        b.close()


def test_six_parameter_xillvercp(cp6):
    """xillverCp (the inclination interpolated as an inner factor, k_xill1 over
    a two-bin angle distribution) on the table: 6 parameter sets with logN
    and Gamma between nodes, FP64; and lmodxillverdensnthcomp."""
    _, x6 = cp6
    rng = np.random.default_rng(63)
    e = log_grid(800, 0.2, 80.0)
    # xillverCp: gamma Afe kTe logxi logN z Incl refl_frac
    ps = [[rng.uniform(1.2, 3.4), rng.uniform(0.5, 10.0), float(np.exp(rng.uniform(np.log(1.0), np.log(400.0)))),
           rng.uniform(0.0, 4.7), rng.uniform(15.0, 19.0), rng.choice([0.0, rng.uniform(0, 1)]),
           rng.uniform(5, 85), rng.choice([-1.0, 0.7])] for _ in range(6)]
    to_oracle = lambda p: p
    b = rx.Batch([dict(model=rx.MODEL_XILLVERCP, energy=e, params=p) for p in ps])
    try:
        outs = _run(b, len(ps), rx.FP64)
    finally:
        b.close()
    for i, p in enumerate(ps):
        ref, st = oracle.xillver_model(oracle.MOD_XILLVERCP, x6, to_oracle(p), e)
        assert st == 0 and outs[i][1] == 0
        err = rel_err(outs[i][0], ref)
        print(f"xillverCp {i}: {err:.3e}", flush=True)
        assert err < TOL64
    y = rx.lmodxillverdensnthcomp(e, ps[0])
    ref, _ = oracle.xillver_model(oracle.MOD_XILLVERCP, x6, to_oracle(ps[0]), e)
    assert rx.lib.rx_last_status() == 0 and rel_err(y, ref) < TOL64


def test_six_parameter_entry_points(cp6):
    """lmodrelxilldensnthcomp and lmodrelxilllpdensnthcomp on the table give
    the FP64 result."""
    T, x6 = cp6
    e = log_grid(700, 0.2, 80.0)
    p = [40.0, 0.7, -1.0, 400.0, 15.0, 3.0, 2.5, 0.0, 2.1, 2.7, 17.3, 3.0, 40.0, 1.5]
    ref, _, st = oracle.relxill_model(oracle.MOD_RELXILLCP, T, x6, p, e)
    assert st == 0
    y = rx.lmodrelxilldensnthcomp(e, p)
    assert rx.lib.rx_last_status() == 0 and conv_err(y, ref) < TOL64
    q = [30.0, 0.9, -1.0, 400.0, 6.0, 0.0, 2.0, 3.1, 18.2, 1.0, 60.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0]
    ref, _, st = oracle.relxill_model(oracle.MOD_RELXILLLPCP, T, x6, q, e)
    assert st == 0
    y = rx.lmodrelxilllpdensnthcomp(e, q)
    print(f"lmodrelxilllpdensnthcomp six-parameter: {conv_err(y, ref):.3e}", flush=True)
    assert rx.lib.rx_last_status() == 0 and conv_err(y, ref) < TOL64
