"""GPU parity: HIP path (through the C ABI) against the FP64 CPU oracle.

Bars: FP64 path within 1e-8 relative per bin, FP32 path within 1e-4, over
bins above 1e-10 of the maximum (the reference's comparison filter,
test/refdata/test_refdata_relxill.sl:162-163).
"""
import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import (RELLINE_CFG0, abs_err, draw_lp_params, draw_relconv_params, log_grid,
                           reflection_spectrum, rel_err, tables)

pytestmark = pytest.mark.gpu

TOL64 = 1e-8
TOL32 = 1e-4


def _report(name, y, yref):
    print(f"{name}: rel {rel_err(y, yref):.3e} abs {abs_err(y, yref):.3e} sum {np.sum(y):.12f} ref {np.sum(yref):.12f}",
          flush=True)


@pytest.fixture(scope="module")
def tabs():
    rel, lp, T = tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    return T


def test_relline_cfg0_fp64(tabs):
    e = log_grid(3000, 0.1, 10.0)
    ref, st = oracle.eval_model(oracle.MOD_RELLINE, tabs, RELLINE_CFG0, e)
    assert st == 0
    y, st = rx.eval_model(rx.MODEL_RELLINE, RELLINE_CFG0, e, precision=rx.FP64)
    _report("relline cfg0 fp64", y, ref)
    assert rel_err(y, ref) < TOL64


def test_relline_cfg0_fp32(tabs):
    e = log_grid(3000, 0.1, 10.0)
    ref, _ = oracle.eval_model(oracle.MOD_RELLINE, tabs, RELLINE_CFG0, e)
    y, st = rx.eval_model(rx.MODEL_RELLINE, RELLINE_CFG0, e, precision=rx.FP32)
    _report("relline cfg0 fp32", y, ref)
    assert rel_err(y, ref) < TOL32


def test_lmodrelline_symbol_matches(tabs):
    e = log_grid(3000, 0.1, 10.0)
    y1 = rx.lmodrelline(e, RELLINE_CFG0)
    y2, _ = rx.eval_model(rx.MODEL_RELLINE, RELLINE_CFG0, e)
    # equal up to the FP64 order of the atomic per-chunk sums
    assert rel_err(y1, y2) < 1e-13


@pytest.mark.parametrize("limb", [1, 2])
def test_relline_limb(tabs, limb):
    e = log_grid(1500, 0.1, 10.0)
    p = list(RELLINE_CFG0)
    p[9] = limb
    p[4] = 0.5
    p[5] = 60.0
    ref, st = oracle.eval_model(oracle.MOD_RELLINE, tabs, p, e)
    assert st == 0
    y, _ = rx.eval_model(rx.MODEL_RELLINE, p, e, precision=rx.FP64)
    _report(f"relline limb{limb} fp64", y, ref)
    assert rel_err(y, ref) < TOL64
    y32, _ = rx.eval_model(rx.MODEL_RELLINE, p, e, precision=rx.FP32)
    assert rel_err(y32, ref) < TOL32


def test_relline_lp_fp64_fp32(tabs):
    e = log_grid(3000, 0.1, 10.0)
    for p in draw_lp_params(6, seed=11):
        ref, st = oracle.eval_model(oracle.MOD_RELLINELP, tabs, p, e)
        assert st == 0, p
        y, _ = rx.eval_model(rx.MODEL_RELLINELP, p, e, precision=rx.FP64)
        _report(f"relline_lp {np.round(p[1:5], 3)} fp64", y, ref)
        assert rel_err(y, ref) < TOL64
        y32, _ = rx.eval_model(rx.MODEL_RELLINELP, p, e, precision=rx.FP32)
        assert rel_err(y32, ref) < TOL32


def test_relconv_fp64(tabs):
    e_ref, refl = reflection_spectrum(n_incl=1)
    e = log_grid(3000, 0.1, 100.0)
    fin = oracle.rebin(e, e_ref, refl[0])
    p = [3.0, 3.0, 15.0, 0.998, 30.0, -1.0, 400.0, 0.0]
    ref, st = oracle.eval_model(oracle.MOD_RELCONV, tabs, p, e, flux_in=fin)
    assert st == 0
    y, _ = rx.eval_model(rx.MODEL_RELCONV, p, e, flux_in=fin, precision=rx.FP64)
    _report("relconv fp64", y, ref)
    assert rel_err(y, ref) < TOL64
    y32, _ = rx.eval_model(rx.MODEL_RELCONV, p, e, flux_in=fin, precision=rx.FP32)
    _report("relconv fp32", y32, ref)
    assert rel_err(y32, ref) < TOL32


def test_relconv_lp_fp64(tabs):
    e_ref, refl = reflection_spectrum(n_incl=1)
    e = log_grid(2000, 0.1, 100.0)
    fin = oracle.rebin(e, e_ref, refl[0])
    p = [10.0, 0.3, 0.7, 45.0, -2.0, 300.0, 0.0, 2.2, 0.0]
    ref, st = oracle.eval_model(oracle.MOD_RELCONVLP, tabs, p, e, flux_in=fin)
    assert st == 0
    y, _ = rx.eval_model(rx.MODEL_RELCONVLP, p, e, flux_in=fin, precision=rx.FP64)
    _report("relconv_lp fp64", y, ref)
    assert rel_err(y, ref) < TOL64


def test_batch_uneven_relconv_fp32(tabs):
    rng = np.random.default_rng(7)
    e_ref, refl = reflection_spectrum(n_incl=1)
    reqs, refs = [], []
    for p in draw_relconv_params(12, seed=9):
        n = int(np.exp(rng.uniform(np.log(300), np.log(8000))))
        e = log_grid(n, 0.1, 100.0)
        fin = oracle.rebin(e, e_ref, refl[0])
        reqs.append(dict(model=rx.MODEL_RELCONV, energy=e, params=p, flux_in=fin))
        ref, st = oracle.eval_model(oracle.MOD_RELCONV, tabs, p, e, flux_in=fin)
        assert st == 0
        refs.append(ref)
    b = rx.Batch(reqs)
    for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
        b.run(prec)
        b.sync()
        for i, ref in enumerate(refs):
            y, st = b.output(i)
            assert st == 0
            assert rel_err(y, ref) < tol, (prec, i, rel_err(y, ref))
    b.close()


def test_fp32_guard_fixup_path(tabs):
    """Widen the FP32 guard band so that nearly every Romberg stop test is
    re-done in FP64 by the fix-up pass: the FP32 path then agrees with the
    oracle at close to FP64 accuracy."""
    e = log_grid(3000, 0.1, 10.0)
    ref, _ = oracle.eval_model(oracle.MOD_RELLINE, tabs, RELLINE_CFG0, e)
    y_default, _ = rx.eval_model(rx.MODEL_RELLINE, RELLINE_CFG0, e, precision=rx.FP32)
    assert rx.lib.rx_set_fp32_guard(0.95) == 0
    try:
        y, _ = rx.eval_model(rx.MODEL_RELLINE, RELLINE_CFG0, e, precision=rx.FP32)
    finally:
        rx.lib.rx_set_fp32_guard(5e-4)
    _report("relline fp32 wide guard", y, ref)
    assert rel_err(y, ref) < TOL32
    assert rel_err(y, ref) <= rel_err(y_default, ref) * 1.5


@pytest.mark.parametrize("n_conv", [1024, 2048, 8192])
def test_relconv_other_conv_grids(tabs, n_conv):
    """Other convolution grid sizes take other FFT plans (radix 8/4/2 passes)
    and workgroup sizes (256 to 1024 threads)."""
    e_ref, refl = reflection_spectrum(n_incl=1)
    e = log_grid(1500, 0.1, 100.0)
    fin = oracle.rebin(e, e_ref, refl[0])
    cg = oracle.log_grid(n_conv + 1, 0.00035, 2000.0)
    p = [3.0, 3.0, 15.0, 0.9, 45.0, -1.0, 400.0, 0.0]
    ref, st = oracle.eval_model(oracle.MOD_RELCONV, tabs, p, e, flux_in=fin, conv_grid=cg)
    assert st == 0
    rx.set_conv_grid(n_conv, 0.00035, 2000.0)
    try:
        for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
            y, st = rx.eval_model(rx.MODEL_RELCONV, p, e, flux_in=fin, precision=prec)
            _report(f"relconv n_conv={n_conv} prec {prec}", y, ref)
            assert st == 0 and rel_err(y, ref) < tol
    finally:
        rx.set_conv_grid(4096, 0.00035, 2000.0)


def test_reference_table_dims(tabs):
    """The released tables' dimensions (src/common.h:69-81): relline 25 spins x
    30 mu0 x 100 radii x 40 g*, lamp post 20 spins x 250 heights x 100 radii
    (40 g* doubles the lane rows in LDS).  relline, relconv and relline_lp
    against the oracle on synthetic tables of those dims, both precisions."""
    rx.tables_synthetic(11, 25, 30, 100, 40)
    rx.lptable_synthetic(13, 20, 250, 100)
    try:
        rel, lp = rx.tables_view()
        T = oracle.Tables({k: (np.array(v) if hasattr(v, "shape") else v) for k, v in rel.items()},
                          {k: (np.array(v) if hasattr(v, "shape") else v) for k, v in lp.items()})
        e = log_grid(3000, 0.1, 10.0)
        e_ref, refl = reflection_spectrum(n_incl=1)
        e_c = log_grid(2000, 0.1, 100.0)
        fin = oracle.rebin(e_c, e_ref, refl[0])
        cases = [(rx.MODEL_RELLINE, oracle.MOD_RELLINE, RELLINE_CFG0, e, None),
                 (rx.MODEL_RELLINE, oracle.MOD_RELLINE, [6.4, 3.0, 3.0, 15.0, 0.3, 75.0, -1.0, 400.0, 0.0, 2.0], e, None),
                 (rx.MODEL_RELCONV, oracle.MOD_RELCONV, [3.0, 3.0, 15.0, -0.4, 45.0, -1.0, 400.0, 0.0], e_c, fin),
                 (rx.MODEL_RELLINELP, oracle.MOD_RELLINELP, [6.4, 8.0, 0.9, 30.0, -1.0, 400.0, 0.0, 0.0, 2.0, 0.0],
                  e, None)]
        for model, omodel, p, grid, flux_in in cases:
            ref, st = oracle.eval_model(omodel, T, p, grid, flux_in=flux_in)
            assert st == 0
            for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
                y, st = rx.eval_model(model, p, grid, flux_in=flux_in, precision=prec)
                _report(f"reference dims model {model} prec {prec}", y, ref)
                assert st == 0 and rel_err(y, ref) < tol
    finally:
        rx.tables_synthetic(1, 25, 22, 100, 20)
        rx.lptable_synthetic(3, 20, 100, 100)


@pytest.mark.parametrize("model", [rx.MODEL_RELLINE, rx.MODEL_RELCONV, rx.MODEL_RELLINELP, rx.MODEL_RELCONVLP])
def test_random_parameter_sweep(tabs, model):
    """48 parameter sets over the full lmodel.dat ranges (limb laws, negative
    radii and heights as multiples of the ISCO / horizon, emissivity indices
    -2..6, rejected sets), each on its own log grid of 300-3000 bins, in one
    batch: accepted where the oracle accepts (same status class otherwise), and
    within the FP64 / FP32 bars where both evaluate."""
    from tests.helpers import draw_wide_params
    rng = np.random.default_rng(1000 + model)
    conv = model in (rx.MODEL_RELCONV, rx.MODEL_RELCONVLP)
    e_ref, refl = reflection_spectrum(n_incl=1)
    reqs, jobs = [], []
    for _ in range(48):
        p = draw_wide_params(model, rng)
        n = int(np.exp(rng.uniform(np.log(300), np.log(3000))))
        e = log_grid(n, 0.1, 100.0 if conv else 10.0)
        fin = oracle.rebin(e, e_ref, refl[0]) if conv else None
        reqs.append(dict(model=model, energy=e, params=p, flux_in=fin))
        jobs.append((model, p, e, fin))
    refs, sts = oracle.eval_batch_mt(tabs, jobs, 16)
    cls = lambda s: 10 if 10 <= s < 50 else s
    b = rx.Batch(reqs)
    n_eval = 0
    worst = {rx.FP64: 0.0, rx.FP32: 0.0}
    try:
        for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
            b.run(prec)
            b.sync()
            for i, r in enumerate(reqs):
                y, st = b.output(i)
                assert cls(st) == cls(sts[i]), (i, r["params"], st, sts[i])
                if st == 0:
                    err = rel_err(y, refs[i])
                    worst[prec] = max(worst[prec], err)
                    assert err < tol, (prec, i, r["params"], err)
                    n_eval += prec == rx.FP64
    finally:
        b.close()
    print(f"sweep model {model}: {n_eval} evaluated, worst rel err FP64 {worst[rx.FP64]:.3e} "
          f"FP32 {worst[rx.FP32]:.3e}", flush=True)
    assert n_eval >= 20
