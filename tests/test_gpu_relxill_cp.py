"""relxillCp and relxilllpCp on the GPU (the nthcomp solves of k_nth, the
nthcomp primary of k_primary, the nthcomp models' xillver table, the
zones' nthcomp normalisation changes and returning-radiation corrections)
against the oracle: lmodel.dat defaults, randomized sweeps over the
lmodel.dat ranges, batches mixing both primaries, the XSPEC entry points,
logN on a table without a density axis, the unsupported ionisation
gradients and the lamp post's zone records.  Spectra are compared with
conv_err: low kTe puts bins 1e-7 to 1e-10 below the maximum inside 0.1-100
keV, where the FP64 round-off of the convolutions (an absolute ~1e-14 to
1e-13 of the maximum, different in the oracle's FFT and the device's) is
1e-7 to 1e-3 of the bin."""
import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import conv_err, draw_relxill_params, log_grid, rel_err, tables

pytestmark = pytest.mark.gpu

TOL64 = 1e-8
TOL32 = 1e-4


@pytest.fixture(scope="module")
def cp():
    _, _, T = tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    rx.xilltable_synthetic(5, 7, 4, 6, 5, 10, 2999)
    rx.xillcptable_synthetic(6, 7, 4, 6, 5, 10, 2999)
    xt = oracle.Xill(rx.xilltable_view())
    xc = oracle.Xill(rx.xilltable_view(nthcomp=True))
    yield T, xt, xc
    rx.xilltable_synthetic()
    rx.xillcptable_synthetic()


def _log_uniform(rng, lo, hi):
    return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))


def draw_cp(rng):
    """relxillCp: Incl a Rin Rout Rbr Index1 Index2 z gamma logxi logN Afe kTe refl_frac"""
    return [rng.uniform(5, 85), rng.uniform(-0.998, 0.998), rng.choice([-1.0, rng.uniform(1.0, 30.0)]),
            rng.uniform(50, 1000), rng.uniform(2, 100), rng.uniform(-1, 6), rng.uniform(-1, 6),
            rng.choice([0.0, rng.uniform(0, 2)]), rng.uniform(1.2, 3.4), rng.uniform(0.0, 4.7),
            rng.choice([15.0, rng.uniform(15.0, 20.0)]), rng.uniform(0.5, 10.0), _log_uniform(rng, 1.0, 400.0),
            rng.choice([-1.0, rng.uniform(-5, 10)])]


def draw_lpcp(rng, returnrad=None):
    """relxilllpCp: Incl a Rin Rout h beta gamma logxi logN Afe kTe refl_frac z
    iongrad_index iongrad_type switch_returnrad switch_reflfrac_boost"""
    rr = float(rng.integers(0, 2)) if returnrad is None else returnrad
    return [rng.uniform(5, 85), rng.uniform(-0.998, 0.998), rng.choice([-1.0, rng.uniform(1.0, 20.0)]),
            rng.uniform(50, 1000), rng.uniform(3.0, 100.0), rng.choice([0.0, rng.uniform(0.0, 0.7)]),
            rng.uniform(1.2, 3.4), rng.uniform(0.0, 4.7), rng.choice([15.0, rng.uniform(15.0, 20.0)]),
            rng.uniform(0.5, 10.0), _log_uniform(rng, 1.0, 400.0), rng.choice([-1.0, rng.uniform(-5, 10)]),
            rng.choice([0.0, rng.uniform(0, 2)]), 0.0, 0.0, rr, float(rng.integers(0, 2))]


def _run(reqs, precision):
    b = rx.Batch(reqs)
    try:
        b.run(precision)
        b.sync()
        return [b.output(i) for i in range(len(reqs))]
    finally:
        b.close()


@pytest.mark.parametrize("model", [rx.MODEL_RELXILLCP, rx.MODEL_RELXILLLPCP])
def test_default_parameters(cp, model):
    """lmodel.dat defaults on 3000 bins over 0.1-100 keV (relxilllpCp with
    switch_returnrad 1 at a = 0.998: the zones' returning-radiation
    corrections use the nthcomp spectra), FP64 and FP32; the angle
    distributions too."""
    T, _, xc = cp
    p = list(rx.DEFAULTS[model])
    om = oracle.MOD_RELXILLCP if model == rx.MODEL_RELXILLCP else oracle.MOD_RELXILLLPCP
    e = log_grid(3000, 0.1, 100.0)
    ref, dist_ref, st = oracle.relxill_model(om, T, xc, p, e)
    assert st == 0
    nd = dist_ref.size
    for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
        b = rx.Batch([dict(model=model, energy=e, params=p)])
        b.run(prec)
        b.sync()
        y, st = b.output(0)
        d = b.dist(0, nd)
        b.close()
        print(f"model {model} defaults prec {prec}: rel {conv_err(y, ref):.3e} dist "
              f"{rel_err(d, dist_ref.reshape(-1)):.3e}", flush=True)
        assert st == 0
        assert conv_err(y, ref) < tol
        assert rel_err(d, dist_ref.reshape(-1)) < tol


@pytest.mark.parametrize("model", [rx.MODEL_RELXILLCP, rx.MODEL_RELXILLLPCP])
def test_random_sweep(cp, model):
    """10 parameter sets over the lmodel.dat ranges (kTe 1-400 keV, logN
    15-20, both signs of refl_frac, redshift; for the lamp post heights 3-100,
    moving sources, returning radiation and the boost switch), one batch,
    FP64 and FP32."""
    T, _, xc = cp
    rng = np.random.default_rng(31 if model == rx.MODEL_RELXILLCP else 32)
    om = oracle.MOD_RELXILLCP if model == rx.MODEL_RELXILLCP else oracle.MOD_RELXILLLPCP
    draw = draw_cp if model == rx.MODEL_RELXILLCP else draw_lpcp
    e = log_grid(1500, 0.1, 100.0)
    ps = [draw(rng) for _ in range(10)]
    refs = [oracle.relxill_model(om, T, xc, p, e) for p in ps]
    assert sum(r[2] == 0 for r in refs) >= 6
    b = rx.Batch([dict(model=model, energy=e, params=p) for p in ps])
    try:
        for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
            b.run(prec)
            b.sync()
            worst = 0.0
            for i, (ref, _, st) in enumerate(refs):
                y, st_g = b.output(i)
                assert (st_g == 0) == (st == 0), (i, ps[i], st, st_g)
                if st == 0:
                    worst = max(worst, conv_err(y, ref))
            print(f"model {model} sweep prec {prec}: worst rel {worst:.3e}", flush=True)
            assert worst < tol
    finally:
        b.close()


def test_both_primaries_in_one_batch(cp):
    """relxill, relxillCp, relxilllp, relxilllpCp and relconv requests in one
    batch, each against the oracle with its own table (FP64)."""
    T, xt, xc = cp
    rng = np.random.default_rng(33)
    e = log_grid(900, 0.1, 100.0)
    flux_in = np.diff(e) * (0.5 * (e[1:] + e[:-1])) ** -2.0
    reqs, refs = [], []
    for _ in range(2):
        p = draw_relxill_params(rng)
        p[12] = 1.5
        reqs.append(dict(model=rx.MODEL_RELXILL, energy=e, params=p))
        refs.append(oracle.relxill_model(oracle.MOD_RELXILL, T, xt, p, e))
        p = draw_cp(rng)
        p[13] = 2.0
        reqs.append(dict(model=rx.MODEL_RELXILLCP, energy=e, params=p))
        refs.append(oracle.relxill_model(oracle.MOD_RELXILLCP, T, xc, p, e))
        p = list(rx.DEFAULTS[rx.MODEL_RELXILLLP])
        p[2], p[10] = rng.uniform(0.0, 0.998), rng.uniform(20, 500)
        reqs.append(dict(model=rx.MODEL_RELXILLLP, energy=e, params=p))
        refs.append(oracle.relxill_model(oracle.MOD_RELXILLLP, T, xt, p, e))
        p = draw_lpcp(rng, returnrad=1.0)
        reqs.append(dict(model=rx.MODEL_RELXILLLPCP, energy=e, params=p))
        refs.append(oracle.relxill_model(oracle.MOD_RELXILLLPCP, T, xc, p, e))
        reqs.append(dict(model=rx.MODEL_RELCONV, energy=e, params=rx.DEFAULTS[rx.MODEL_RELCONV], flux_in=flux_in))
        refs.append(None)
    out = _run(reqs, rx.FP64)
    for i, ref in enumerate(refs):
        if ref is None:
            continue
        y, st = out[i]
        assert (st == 0) == (ref[2] == 0), (i, reqs[i]["model"], st, ref[2])
        if st == 0:
            err = conv_err(y, ref[0])
            print(f"request {i} model {reqs[i]['model']}: rel {err:.3e}", flush=True)
            assert err < TOL64


def test_entry_points(cp):
    """lmodrelxilldensnthcomp and lmodrelxilllpdensnthcomp give the FP64
    result; rejected parameters give zeros and a status."""
    T, _, xc = cp
    e = log_grid(700, 0.2, 80.0)
    p = [40.0, 0.7, -1.0, 400.0, 20.0, 3.0, 2.5, 0.05, 2.2, 1.5, 16.0, 3.0, 30.0, 2.0]
    ref, _, st = oracle.relxill_model(oracle.MOD_RELXILLCP, T, xc, p, e)
    assert st == 0
    y = rx.lmodrelxilldensnthcomp(e, p)
    assert rx.lib.rx_last_status() == 0 and conv_err(y, ref) < TOL64
    q = [50.0, 0.7, -1.0, 300.0, 10.0, 0.1, 2.3, 2.0, 15.0, 2.0, 150.0, 2.0, 0.02, 0.0, 0.0, 1.0, 0.0]
    ref, _, st = oracle.relxill_model(oracle.MOD_RELXILLLPCP, T, xc, q, e)
    assert st == 0
    y = rx.lmodrelxilllpdensnthcomp(e, q)
    print(f"lmodrelxilllpdensnthcomp: rel {conv_err(y, ref):.3e}", flush=True)
    assert rx.lib.rx_last_status() == 0 and conv_err(y, ref) < TOL64
    bad = list(p)
    bad[2], bad[3] = 50.0, 20.0  # Rin > Rout
    y = rx.lmodrelxilldensnthcomp(e, bad)
    assert not y.any() and rx.lib.rx_last_status() == rx.RX_ERR_PARAM
    grad = list(q)
    grad[14] = 2.0  # the alpha-disc ionisation gradient (not built)
    y = rx.lmodrelxilllpdensnthcomp(e, grad)
    assert not y.any() and rx.lib.rx_last_status() == rx.RX_ERR_UNSUPPORTED


def test_density_on_a_table_without_density_axis(cp):
    """logN != 15 on a five-parameter table divides its rows by
    10^(logN - 15) in float, as the reference's loader does: FP64 parity for
    logN 15, 16.3 and 19.5, reflection only and with the primary."""
    T, _, xc = cp
    e = log_grid(1000, 0.1, 100.0)
    ps = []
    for dens in (15.0, 16.3, 19.5):
        for rf in (-1.0, 2.0):
            p = list(rx.DEFAULTS[rx.MODEL_RELXILLCP])
            p[10], p[13] = dens, rf
            ps.append(p)
    out = _run([dict(model=rx.MODEL_RELXILLCP, energy=e, params=p) for p in ps], rx.FP64)
    for p, (y, st) in zip(ps, out):
        ref, _, st_o = oracle.relxill_model(oracle.MOD_RELXILLCP, T, xc, p, e)
        assert st == 0 and st_o == 0
        assert conv_err(y, ref) < TOL64, (p[10], p[13])


def test_relxilllpcp_zone_records(cp):
    """rx_batch_zones of relxilllpCp against the oracle's zones: emission
    angle, energy shift and the nthcomp normalisation ratio at 1e-12; the
    zone's kTe (the source's times its shift) bracketed by the kTe axis."""
    T, _, xc = cp
    xv = rx.xilltable_view(nthcomp=True)
    ax = np.asarray(xv["val"][xv["pid"].index(3)], dtype=np.float64)
    rng = np.random.default_rng(34)
    ps = [draw_lpcp(rng, returnrad=0.0) for _ in range(6)]
    ps = [p for p in ps if oracle.relxill_zones(oracle.MOD_RELXILLLPCP, T, p)[1] == 0]
    assert len(ps) >= 3
    b = rx.Batch([dict(model=rx.MODEL_RELXILLLPCP, energy=log_grid(500, 0.1, 100.0), params=p) for p in ps])
    try:
        b.run(rx.FP64)
        b.sync()
        for i, p in enumerate(ps):
            zo, st = oracle.relxill_zones(oracle.MOD_RELXILLLPCP, T, p)
            zg = b.zones(i, 10)
            np.testing.assert_allclose(zg[:, 0], 0.5 * (zo[:, 0] + zo[:, 1]), rtol=1e-12)
            np.testing.assert_allclose(zg[:, 1:4], zo[:, [2, 3, 5]], rtol=1e-12, atol=1e-300)
            for z in range(10):
                k = int(zg[z, 4])
                kte = min(max(zo[z, 4], ax[0]), ax[-1])
                assert 0 <= k < ax.size - 1 and ax[k] <= kte * (1 + 1e-7) and kte <= ax[k + 1] * (1 + 1e-7)
    finally:
        b.close()
