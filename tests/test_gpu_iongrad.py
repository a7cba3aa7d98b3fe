"""relxilllpCp's power-law ionisation gradient on the GPU (25 zones, each with
its own logXi bracket from zone_record; one k_xill1 unit per zone; k_setup
with 25-zone arrays) against the oracle: a randomized sweep over the
lmodel.dat ranges (with and without returning radiation), the zones' logXi
brackets, a batch mixing gradient, constant-ionisation and relxill requests,
and the XSPEC entry point.  Spectra are compared with conv_err."""
import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import conv_err, draw_relxill_params, log_grid, rel_err, tables
from tests.test_gpu_relxill_cp import draw_lpcp

pytestmark = pytest.mark.gpu

TOL64 = 1e-8
TOL32 = 1e-4


@pytest.fixture(scope="module")
def cpg():
    _, _, T = tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    rx.xilltable_synthetic(5, 7, 4, 6, 5, 10, 2999)
    rx.xillcptable_synthetic(6, 7, 4, 6, 5, 10, 2999)
    x = rx.xilltable_view(nthcomp=True)
    yield T, oracle.Xill(rx.xilltable_view()), oracle.Xill(x), x
    rx.xilltable_synthetic()
    rx.xillcptable_synthetic()


def draw_grad(rng, returnrad):
    p = draw_lpcp(rng, returnrad=returnrad)
    p[13] = rng.uniform(-2.0, 3.0)  # iongrad_index
    p[14] = 1.0                     # iongrad_type: the power law
    return p


def test_gradient_sweep(cpg):
    """8 gradient requests (index -2..3, logXi 0-4.7, half with returning
    radiation), one batch, FP64 and FP32."""
    T, _, xc, _ = cpg
    rng = np.random.default_rng(71)
    e = log_grid(1200, 0.1, 100.0)
    ps = [draw_grad(rng, float(k % 2)) for k in range(8)]
    refs = [oracle.relxill_model(oracle.MOD_RELXILLLPCP, T, xc, p, e) for p in ps]
    assert sum(r[2] == 0 for r in refs) >= 5
    b = rx.Batch([dict(model=rx.MODEL_RELXILLLPCP, energy=e, params=p) for p in ps])
    try:
        for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
            b.run(prec)
            b.sync()
            worst = 0.0
            for i, (ref, dist_ref, st) in enumerate(refs):
                y, st_g = b.output(i)
                assert (st_g == 0) == (st == 0), (i, ps[i], st, st_g)
                if st == 0:
                    worst = max(worst, conv_err(y, ref))
                    assert rel_err(b.dist(i, dist_ref.size), dist_ref.reshape(-1)) < 1e-8
            print(f"ionisation gradient prec {prec}: worst {worst:.3e}", flush=True)
            assert worst < tol
    finally:
        b.close()


def test_gradient_zone_brackets(cpg):
    """Every zone's logXi bracket (rx_batch_zones columns 6, 7) is the
    oracle's zone logXi bracketed on the table's axis: index and float factor."""
    T, _, _, x = cpg
    ax = np.asarray(x["val"][2], dtype=np.float32)
    rng = np.random.default_rng(72)
    ps = [draw_grad(rng, 0.0) for _ in range(6)]
    ps = [p for p in ps if oracle.relxill_zones(oracle.MOD_RELXILLLPCP, T, p)[1] == 0]
    assert len(ps) >= 3
    b = rx.Batch([dict(model=rx.MODEL_RELXILLLPCP, energy=log_grid(300, 0.1, 100.0), params=p) for p in ps])
    try:
        b.run(rx.FP64)
        b.sync()
        for i, p in enumerate(ps):
            zo, _ = oracle.relxill_zones(oracle.MOD_RELXILLLPCP, T, p)
            zg = b.zones(i, 25)
            np.testing.assert_allclose(zg[:, 0], 0.5 * (zo[:, 0] + zo[:, 1]), rtol=1e-12)
            for z in range(25):
                v = np.float32(min(max(np.float32(zo[z, 6]), ax[0]), ax[-1]))
                k = int(zg[z, 6])
                assert 0 <= k < ax.size - 1 and ax[k] <= v <= ax[k + 1], (i, z)
                assert abs(zg[z, 7] - float(np.float32((v - ax[k]) / (ax[k + 1] - ax[k])))) < 1e-7, (i, z)
    finally:
        b.close()


def test_mixed_batch(cpg):
    """Gradient, constant-ionisation relxilllpCp and relxill requests in one
    batch (k_setup with the 25-zone arrays for all of them), FP64."""
    T, xt, xc, _ = cpg
    rng = np.random.default_rng(73)
    e = log_grid(900, 0.1, 100.0)
    reqs, refs = [], []
    for k in range(2):
        p = draw_grad(rng, float(k))
        reqs.append(dict(model=rx.MODEL_RELXILLLPCP, energy=e, params=p))
        refs.append(oracle.relxill_model(oracle.MOD_RELXILLLPCP, T, xc, p, e))
        p = draw_lpcp(rng, returnrad=float(k))
        reqs.append(dict(model=rx.MODEL_RELXILLLPCP, energy=e, params=p))
        refs.append(oracle.relxill_model(oracle.MOD_RELXILLLPCP, T, xc, p, e))
        p = draw_relxill_params(rng)
        reqs.append(dict(model=rx.MODEL_RELXILL, energy=e, params=p))
        refs.append(oracle.relxill_model(oracle.MOD_RELXILL, T, xt, p, e))
    b = rx.Batch(reqs)
    try:
        b.run(rx.FP64)
        b.sync()
        for i, ref in enumerate(refs):
            y, st = b.output(i)
            assert (st == 0) == (ref[2] == 0), (i, st, ref[2])
            if st == 0:
                err = conv_err(y, ref[0])
                print(f"mixed request {i}: {err:.3e}", flush=True)
                assert err < TOL64
    finally:
        b.close()


def test_gradient_entry_point(cpg):
    """lmodrelxilllpdensnthcomp with iongrad_type 1 gives the FP64 result."""
    T, _, xc, _ = cpg
    e = log_grid(700, 0.2, 80.0)
    q = [50.0, 0.7, -1.0, 300.0, 10.0, 0.1, 2.3, 3.4, 15.0, 2.0, 150.0, 2.0, 0.02, 1.2, 1.0, 1.0, 0.0]
    ref, _, st = oracle.relxill_model(oracle.MOD_RELXILLLPCP, T, xc, q, e)
    assert st == 0
    y = rx.lmodrelxilllpdensnthcomp(e, q)
    print(f"lmodrelxilllpdensnthcomp with a gradient: {conv_err(y, ref):.3e}", flush=True)
    assert rx.lib.rx_last_status() == 0 and conv_err(y, ref) < TOL64
