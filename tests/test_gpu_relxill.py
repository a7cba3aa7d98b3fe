"""relxill on the GPU (k_setup angle distribution, k_xill table interpolation
and angle weighting, k_conv convolution with the primary continuum) against
the oracle: the reference's default parameters, a randomized sweep over the
lmodel.dat ranges, parameters outside the table, both signs of refl_frac,
redshift, a table read from an OGIP FITS file, the XSPEC entry point,
batching with the other models and the serving loop."""
import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import (affine_xilltable, draw_relxill_params, log_grid, rel_err, tables,
                           write_fits_xilltable)

pytestmark = pytest.mark.gpu

TOL64 = 1e-8
TOL32 = 1e-4
XDIMS = (5, 7, 4, 6, 5, 10, 2999)  # seed, Gamma, A_Fe, logXi, Ecut, Incl, energy bins


@pytest.fixture(scope="module")
def relx():
    _, _, T = tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    rx.xilltable_synthetic(*XDIMS)
    xt = oracle.Xill(rx.xilltable_view())
    yield T, xt
    rx.xilltable_synthetic()


def _gpu(params, e, precision, dist=False):
    b = rx.Batch([dict(model=rx.MODEL_RELXILL, energy=e, params=params)])
    b.run(precision)
    b.sync()
    y, st = b.output(0)
    d = b.dist(0, 10) if dist else None
    b.close()
    return y, st, d


def test_relxill_default_parameters(relx):
    """lmodel.dat defaults (Index 3/3, Rbr 15, a 0.998, 30 deg, Rin ISCO,
    Rout 400, z 0, Gamma 2, logxi 3.1, Afe 1, Ecut 300, refl_frac 3) on 3000
    bins over 0.1-100 keV; the angle distribution too."""
    T, xt = relx
    p = rx.DEFAULTS[rx.MODEL_RELXILL]
    e = log_grid(3000, 0.1, 100.0)
    ref, dist_ref, st = oracle.relxill(T, xt, p, e)
    assert st == 0
    for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
        y, st, d = _gpu(p, e, prec, dist=True)
        print(f"relxill default prec {prec}: rel {rel_err(y, ref):.3e} dist {rel_err(d, dist_ref):.3e}")
        assert st == 0
        assert rel_err(y, ref) < tol
        assert rel_err(d, dist_ref) < tol


def test_relxill_random_sweep(relx):
    """12 parameter sets over the full ranges (both signs of refl_frac,
    redshift, Rin beyond the ISCO, xillver parameters across the table), one
    batch, FP64 and FP32."""
    T, xt = relx
    rng = np.random.default_rng(11)
    e = log_grid(1500, 0.1, 100.0)
    ps = [draw_relxill_params(rng) for _ in range(12)]
    refs = []
    for p in ps:
        ref, _, st = oracle.relxill(T, xt, p, e)
        refs.append((ref, st))
    b = rx.Batch([dict(model=rx.MODEL_RELXILL, energy=e, params=p) for p in ps])
    for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
        b.run(prec)
        b.sync()
        worst = 0.0
        for i, (ref, st) in enumerate(refs):
            y, st_g = b.output(i)
            assert st_g == (0 if st == 0 else rx.RX_ERR_PARAM), (i, ps[i], st)
            if st == 0:
                worst = max(worst, rel_err(y, ref))
        print(f"relxill sweep prec {prec}: worst rel {worst:.3e}")
        assert worst < tol
    b.close()


@pytest.mark.parametrize("xpar", [(0.5, 0.1, -2.0, 2.0), (4.0, 30.0, 9.0, 5000.0), (1.0, 0.5, 0.0, 5.0)])
def test_relxill_xillver_parameters_outside_the_table(relx, xpar):
    """Gamma, A_Fe, logXi, Ecut below / above / on the table ends: the values
    are clamped (resetInpvalsToBoundaries, logXi within 0-4.7), Ecut's factor
    pinned to 0 / 1."""
    T, xt = relx
    gam, afe, lxi, ect = xpar
    p = [3.0, 3.0, 15.0, 0.9, 45.0, -1.0, 400.0, 0.0, gam, lxi, afe, ect, 1.0]
    e = log_grid(800, 0.1, 100.0)
    ref, _, st = oracle.relxill(T, xt, p, e)
    assert st == 0
    y, st, _ = _gpu(p, e, rx.FP64)
    assert st == 0
    assert rel_err(y, ref) < TOL64


@pytest.mark.parametrize("rf,z", [(-1.0, 0.0), (0.0, 0.0), (3.0, 0.0), (-2.5, 0.7), (1.0, 2.0)])
def test_relxill_refl_frac_and_redshift(relx, rf, z):
    T, xt = relx
    p = [3.0, 3.0, 15.0, 0.5, 60.0, -1.0, 400.0, z, 1.9, 2.7, 2.0, 150.0, rf]
    e = log_grid(1000, 0.1, 100.0)
    ref, _, st = oracle.relxill(T, xt, p, e)
    assert st == 0
    for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
        y, st, _ = _gpu(p, e, prec)
        assert st == 0
        assert rel_err(y, ref) < tol, (rf, z, prec)


def test_relxill_table_from_fits(relx, tmp_path):
    """A table written as an OGIP file (xillver-a-Ec5 layout) and loaded with
    rx_xilltable_load_fits: the GPU result equals the oracle on the same
    table."""
    T, _ = relx
    x, _ = affine_xilltable(n_ener=500)
    path = tmp_path / "xillver-a-Ec5.fits"
    write_fits_xilltable(path, x)
    try:
        assert rx.lib.rx_xilltable_load_fits(str(path).encode()) == 0, rx.lib.rx_last_error()
        xt = oracle.Xill(rx.xilltable_view())
        e = log_grid(600, 0.1, 100.0)
        for p in ([3.0, 3.0, 15.0, 0.998, 30.0, -1.0, 400.0, 0.0, 2.0, 3.1, 1.0, 300.0, -1.0],
                  [2.0, 4.0, 8.0, -0.5, 70.0, 3.0, 300.0, 0.1, 2.9, 0.7, 6.0, 40.0, 2.0]):
            ref, _, st = oracle.relxill(T, xt, p, e)
            assert st == 0
            y, st, _ = _gpu(p, e, rx.FP64)
            assert st == 0
            assert rel_err(y, ref) < TOL64
    finally:
        rx.xilltable_synthetic(*XDIMS)


def test_lmodrelxill_entry_point(relx):
    """The XSPEC entry point gives the FP64 result; rejected parameters give
    zeros and a status (Rout <= Rin)."""
    T, xt = relx
    p = [3.0, 2.5, 20.0, 0.7, 40.0, -1.0, 400.0, 0.05, 2.2, 1.5, 3.0, 100.0, 2.0]
    e = log_grid(700, 0.2, 80.0)
    ref, _, st = oracle.relxill(T, xt, p, e)
    y = rx.lmodrelxill(e, p)
    assert rel_err(y, ref) < TOL64
    bad = list(p)
    bad[5], bad[6] = 50.0, 20.0
    y = rx.lmodrelxill(e, bad)
    assert not y.any() and rx.lib.rx_last_status() == rx.RX_ERR_PARAM


def test_relxill_batched_with_other_models(relx):
    """relxill requests batched with relline and relconv requests give what
    they give alone (to the order of the FP64 atomic sums)."""
    T, xt = relx
    e = log_grid(900, 0.1, 100.0)
    rng = np.random.default_rng(3)
    px = [draw_relxill_params(rng) for _ in range(4)]
    px = [p for p in px if oracle.relxill(T, xt, p, e)[2] == 0][:3]
    alone = [_gpu(p, e, rx.FP64)[0] for p in px]
    el = log_grid(500, 0.5, 10.0)
    flux_in = np.diff(e) * (0.5 * (e[1:] + e[:-1])) ** -2.0
    reqs = []
    for p in px:
        reqs.append(dict(model=rx.MODEL_RELLINE, energy=el, params=rx.DEFAULTS[rx.MODEL_RELLINE]))
        reqs.append(dict(model=rx.MODEL_RELXILL, energy=e, params=p))
        reqs.append(dict(model=rx.MODEL_RELCONV, energy=e, params=rx.DEFAULTS[rx.MODEL_RELCONV], flux_in=flux_in))
    b = rx.Batch(reqs)
    rx.lib.rx_profile_enable(1)
    try:
        b.run(rx.FP64)
        b.sync()
        ms = b.kernel_ms()
    finally:
        rx.lib.rx_profile_enable(0)
    for k, y0 in enumerate(alone):
        y, st = b.output(3 * k + 1)
        assert st == 0
        assert rel_err(y, y0) < 1e-12
    b.close()
    # setup, integrate, fixup, finalize, conv, xill, primary
    assert len(ms) == 7 and ms[4] > 0 and ms[5] > 0 and ms[6] > 0


def test_relxill_through_the_server(relx):
    """relxill requests admitted by the serving loop return the batch result."""
    T, xt = relx
    e = log_grid(1200, 0.1, 100.0)
    p = [3.0, 3.0, 15.0, 0.9, 35.0, -1.0, 400.0, 0.0, 2.1, 3.0, 1.0, 200.0, 1.5]
    y0, _, _ = _gpu(p, e, rx.FP32)
    s = rx.Server(max_inflight=8, hbm_budget=1 << 30, n_streams=2, precision=rx.FP32)
    try:
        tickets = [s.submit(rx.MODEL_RELXILL, e, p) for _ in range(5)]
        s.drain()
        for t in tickets:
            y, st, _ = s.poll(t)
            assert st == 0
            assert rel_err(y, y0) < 1e-12
    finally:
        s.stop()


# ---- relxilllp: lamp post, 10 radial zones ----

def draw_relxilllp_params(rng):
    a = rng.uniform(-0.998, 0.998)
    return [rng.uniform(3.0, 100.0), rng.choice([0.0, rng.uniform(0.0, 0.7)]), a, rng.uniform(5, 85),
            rng.choice([-1.0, rng.uniform(1.0, 20.0)]), rng.uniform(50, 1000), rng.choice([0.0, rng.uniform(0, 2)]),
            rng.uniform(1.0, 3.4), rng.uniform(0.0, 4.7), rng.uniform(0.5, 10.0), rng.uniform(5.0, 1000.0),
            rng.choice([-1.0, rng.uniform(-5, 10)]), 0.0, float(rng.integers(0, 2))]


def _gpu_lp(params, e, precision, n_dist=0):
    b = rx.Batch([dict(model=rx.MODEL_RELXILLLP, energy=e, params=params)])
    b.run(precision)
    b.sync()
    y, st = b.output(0)
    d = b.dist(0, n_dist) if n_dist else None
    b.close()
    return y, st, d


def test_relxilllp_default_parameters(relx):
    """lmodel.dat defaults (h 6, beta 0, a 0.998, 30 deg, Rin ISCO, Rout 400,
    Gamma 2, logxi 3.1, Afe 1, Ecut 300, refl_frac 1) without returning
    radiation, on 3000 bins over 0.1-100 keV; the ten zones' angle
    distributions too."""
    T, xt = relx
    p = list(rx.DEFAULTS[rx.MODEL_RELXILLLP])
    p[12] = 0.0
    e = log_grid(3000, 0.1, 100.0)
    ref, dist_ref, st = oracle.relxill_model(oracle.MOD_RELXILLLP, T, xt, p, e)
    assert st == 0
    for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
        y, st, d = _gpu_lp(p, e, prec, n_dist=100)
        print(f"relxilllp default prec {prec}: rel {rel_err(y, ref):.3e} dist {rel_err(d, dist_ref.reshape(-1)):.3e}")
        assert st == 0
        assert rel_err(y, ref) < tol
        assert rel_err(d, dist_ref.reshape(-1)) < tol


def test_relxilllp_random_sweep(relx):
    """12 lamp-post parameter sets over the full ranges (heights 3-100, moving
    sources, both signs of refl_frac, the boost switch, redshift), one batch,
    FP64 and FP32."""
    T, xt = relx
    rng = np.random.default_rng(21)
    e = log_grid(1500, 0.1, 100.0)
    ps = [draw_relxilllp_params(rng) for _ in range(12)]
    refs = [oracle.relxill_model(oracle.MOD_RELXILLLP, T, xt, p, e) for p in ps]
    b = rx.Batch([dict(model=rx.MODEL_RELXILLLP, energy=e, params=p) for p in ps])
    for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
        b.run(prec)
        b.sync()
        worst = 0.0
        for i, (ref, _, st) in enumerate(refs):
            y, st_g = b.output(i)
            assert (st_g == 0) == (st == 0), (i, ps[i], st, st_g)
            if st == 0:
                worst = max(worst, rel_err(y, ref))
        print(f"relxilllp sweep prec {prec}: worst rel {worst:.3e}")
        assert worst < tol
    b.close()


def test_relxilllp_entry_point_and_returning_radiation(relx):
    """lmodrelxilllp gives the FP64 result, at its lmodel.dat defaults too
    (switch_returnrad = 1, a = 0.998: returning radiation with the zones'
    xillver correction factors)."""
    T, xt = relx
    p = [10.0, 0.1, 0.7, 50.0, -1.0, 300.0, 0.02, 2.3, 2.0, 2.0, 150.0, 2.0, 0.0, 1.0]
    e = log_grid(700, 0.2, 80.0)
    ref, _, st = oracle.relxill_model(oracle.MOD_RELXILLLP, T, xt, p, e)
    assert st == 0
    y = rx.lmodrelxilllp(e, p)
    assert rel_err(y, ref) < TOL64
    d = list(rx.DEFAULTS[rx.MODEL_RELXILLLP])
    assert d[12] == 1.0
    ref, _, st = oracle.relxill_model(oracle.MOD_RELXILLLP, T, xt, d, e)
    assert st == 0
    y = rx.lmodrelxilllp(e, d)
    print(f"lmodrelxilllp defaults: rel {rel_err(y, ref):.3e}", flush=True)
    assert rx.lib.rx_last_status() == 0 and rel_err(y, ref) < TOL64


@pytest.mark.parametrize("switch", [1.0, 2.0, -1.0])
def test_relxilllp_returning_radiation(relx, switch):
    """relxilllp with returning radiation: at spins <= 0 the plain returning
    emissivity, above 0 with the zones' xillver correction factors
    (src/Relxill.cpp:337-344); heights, moving sources, both signs of
    refl_frac; one batch, FP64 and FP32, outputs and the zones' angle
    distributions against the oracle."""
    T, xt = relx
    e = log_grid(1200, 0.1, 100.0)
    ps = []
    for a, h, beta, rf in ((-0.6, 6.0, 0.0, 1.0), (0.0, 12.0, 0.2, -1.0), (0.5, 4.0, 0.0, 2.0),
                           (0.998, 6.0, 0.0, 1.0), (0.9, 30.0, 0.4, -2.0), (0.3, 3.5, 0.1, 0.5)):
        ps.append([h, beta, a, 35.0, -1.0, 400.0, 0.0, 2.1, 2.7, 1.5, 120.0, rf, switch, 0.0])
    refs = [oracle.relxill_model(oracle.MOD_RELXILLLP, T, xt, p, e) for p in ps]
    b = rx.Batch([dict(model=rx.MODEL_RELXILLLP, energy=e, params=p) for p in ps])
    try:
        for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
            b.run(prec)
            b.sync()
            worst, worst_d = 0.0, 0.0
            for i, (ref, dist_ref, st) in enumerate(refs):
                assert st == 0, (i, st)
                y, st_g = b.output(i)
                assert st_g == 0, (i, rx.lib.rx_last_error())
                worst = max(worst, rel_err(y, ref))
                worst_d = max(worst_d, rel_err(b.dist(i, 100), dist_ref.reshape(-1)))
            print(f"relxilllp returning radiation switch {switch} prec {prec}: rel {worst:.3e} dist {worst_d:.3e}",
                  flush=True)
            assert worst < tol and worst_d < tol
    finally:
        b.close()


def test_relxilllp_batched_with_relxill(relx):
    """relxilllp requests batched with relxill and relconv requests give what
    they give alone (to the order of the FP64 atomic sums)."""
    T, xt = relx
    e = log_grid(900, 0.1, 100.0)
    rng = np.random.default_rng(8)
    pl = [draw_relxilllp_params(rng) for _ in range(3)]
    alone = [_gpu_lp(p, e, rx.FP64)[0] for p in pl]
    flux_in = np.diff(e) * (0.5 * (e[1:] + e[:-1])) ** -2.0
    reqs = []
    for p in pl:
        reqs.append(dict(model=rx.MODEL_RELXILL, energy=e, params=rx.DEFAULTS[rx.MODEL_RELXILL]))
        reqs.append(dict(model=rx.MODEL_RELXILLLP, energy=e, params=p))
        reqs.append(dict(model=rx.MODEL_RELCONV, energy=e, params=rx.DEFAULTS[rx.MODEL_RELCONV], flux_in=flux_in))
    b = rx.Batch(reqs)
    b.run(rx.FP64)
    b.sync()
    for k, y0 in enumerate(alone):
        y, st = b.output(3 * k + 1)
        assert st == 0
        assert rel_err(y, y0) < 1e-12
    b.close()


def test_relxilllp_zone_records(relx):
    """rx_batch_zones against the oracle's zones: emission angle, energy shift
    source -> zone and normalisation ratio at 1e-12, the zone's Ecut bracketed
    by the table nodes its index names, the factor at FP32 rounding."""
    T, xt = relx
    xv = rx.xilltable_view()
    ax = np.asarray(xv["val"][xv["pid"].index(3)], dtype=np.float64)
    rng = np.random.default_rng(12)
    ps = []
    for _ in range(6):
        a = rng.uniform(-0.998, 0.998)
        ps.append([rng.uniform(3.0, 100.0), rng.uniform(0.0, 0.5), a, rng.uniform(5.0, 85.0), -1.0, 400.0, 0.0,
                   rng.uniform(1.2, 3.2), 3.1, 1.0, float(np.exp(rng.uniform(np.log(8.0), np.log(800.0)))), 1.0,
                   0.0, 0.0])
    b = rx.Batch([dict(model=rx.MODEL_RELXILLLP, energy=log_grid(500, 0.1, 100.0), params=p) for p in ps])
    try:
        b.run(rx.FP64)
        b.sync()
        for i, p in enumerate(ps):
            zo, st = oracle.relxill_zones(oracle.MOD_RELXILLLP, T, p)
            assert st == 0
            zg = b.zones(i, 10)
            np.testing.assert_allclose(zg[:, 0], 0.5 * (zo[:, 0] + zo[:, 1]), rtol=1e-12)
            np.testing.assert_allclose(zg[:, 1:4], zo[:, [2, 3, 5]], rtol=1e-12, atol=1e-300)
            for z in range(10):
                k = int(zg[z, 4])
                ect = min(max(zo[z, 4], ax[0]), ax[-1])
                assert 0 <= k < ax.size - 1 and ax[k] <= ect * (1 + 1e-7) and ect <= ax[k + 1] * (1 + 1e-7)
                assert abs(zg[z, 5] - (ect - ax[k]) / (ax[k + 1] - ax[k])) < 1e-6
        with pytest.raises(rx.RxError):
            b.zones(0, 11)
    finally:
        b.close()
