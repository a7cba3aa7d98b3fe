"""Returning radiation on the GPU (k_setup's returning-emissivity block) for
relline_lp and relconv_lp against the oracle: switch_returnrad 1 (direct plus
returning), 2 / -1 (returning only), a randomized batch, and a table read from
the FITS layout of table_returnRad_v20220301.fits."""
import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import (draw_lp_params, log_grid, reflection_spectrum, rel_err, rrad_table, tables,
                           write_fits_rradtable)

pytestmark = pytest.mark.gpu

TOL64 = 1e-8
TOL32 = 1e-4


@pytest.fixture(scope="module")
def tabs():
    _, _, T = tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    return T


def _check(T, model, omodel, p, e, fin=None, tag=""):
    ref, st = oracle.eval_model(omodel, T, p, e, flux_in=fin)
    assert st == 0, p
    for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
        y, st = rx.eval_model(model, p, e, flux_in=fin, precision=prec)
        err = rel_err(y, ref)
        print(f"rrad {tag} {np.round(p, 3).tolist()} prec {prec}: rel {err:.3e}", flush=True)
        assert st == 0 and err < tol


@pytest.mark.parametrize("switch", [1.0, 2.0, -1.0])
def test_relline_lp_switches(tabs, switch):
    e = log_grid(3000, 0.1, 10.0)
    for p in draw_lp_params(4, seed=21):
        p[9] = switch
        _check(tabs, rx.MODEL_RELLINELP, oracle.MOD_RELLINELP, p, e, tag=f"relline_lp sw{switch}")


def test_relconv_lp_with_returning_radiation(tabs):
    e_ref, refl = reflection_spectrum(n_incl=1)
    e = log_grid(2000, 0.1, 100.0)
    fin = oracle.rebin(e, e_ref, refl[0])
    for p in ([6.0, 0.0, 0.998, 30.0, -1.0, 400.0, 0.0, 2.0, 1.0],
              [10.0, 0.3, 0.7, 45.0, -2.0, 300.0, 0.0, 2.2, 1.0],
              [3.5, 0.0, -0.6, 70.0, 8.0, 900.0, 1.0, 1.6, 2.0]):
        _check(tabs, rx.MODEL_RELCONVLP, oracle.MOD_RELCONVLP, p, e, fin, tag="relconv_lp")


def test_random_batch_with_returning_radiation(tabs):
    """24 relline_lp sets (spins over the range, heights 3-100, Rin beyond the
    ISCO) with switch 1, one batch, each on its own grid."""
    rng = np.random.default_rng(31)
    reqs, refs = [], []
    for p in draw_lp_params(24, seed=32):
        p[9] = 1.0
        p[8] = rng.uniform(1.2, 3.2)
        e = log_grid(int(rng.integers(300, 3000)), 0.1, 10.0)
        ref, st = oracle.eval_model(oracle.MOD_RELLINELP, tabs, p, e)
        assert st == 0
        reqs.append(dict(model=rx.MODEL_RELLINELP, energy=e, params=p))
        refs.append(ref)
    b = rx.Batch(reqs)
    try:
        for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
            b.run(prec)
            b.sync()
            worst = 0.0
            for i, ref in enumerate(refs):
                y, st = b.output(i)
                assert st == 0
                worst = max(worst, rel_err(y, ref))
            print(f"rrad batch prec {prec}: worst {worst:.3e}", flush=True)
            assert worst < tol
    finally:
        b.close()


def test_fits_rradtable(tabs, tmp_path):
    """A table with returning fractions shifted one zone outward, 1.2x boosted
    photons and spin-dependent scale, loaded from the FITS layout."""
    t = rrad_table(n_spin=3, n_rad=20, n_g=8, c=0.04, shift=1, g0=1.2, spin_scale=(1.0, 1.5, 2.0))
    path = tmp_path / "table_returnRad_v20220301.fits"
    write_fits_rradtable(path, t)
    rel, lp, _ = tables()
    try:
        assert rx.lib.rx_rradtable_load_fits(str(path).encode()) == 0, rx.lib.rx_last_error()
        T = oracle.Tables(rel, lp, rx.rradtable_view())
        e = log_grid(2000, 0.1, 10.0)
        for p in ([6.4, 6.0, 0.998, 30.0, -1.0, 400.0, 0.0, 0.0, 2.0, 1.0],
                  [6.4, 12.0, 0.2, 60.0, -1.5, 800.0, 0.0, 1.0, 2.5, 1.0],
                  [6.4, 4.0, -0.7, 20.0, 15.0, 400.0, 0.3, 2.0, 1.8, 2.0]):
            _check(T, rx.MODEL_RELLINELP, oracle.MOD_RELLINELP, p, e, tag="fits")
    finally:
        rx.rradtable_synthetic(7, 21, 50, 20)


def test_rrad_table_limits(tabs, tmp_path):
    """A spin above every tabulated spin takes the last one, whose innermost
    zone starts outside Rin: refused by both; more than 64 table zones on the
    disc is this library's documented limit (status, zero flux), where the
    oracle still evaluates."""
    rel, lp, _ = tables()
    e = log_grid(500, 0.1, 10.0)
    p = [6.4, 6.0, 0.9, 30.0, -1.0, 400.0, 0.0, 0.0, 2.0, 1.0]
    try:
        t = rrad_table(n_spin=2, n_rad=10, n_g=4, spins=(-0.5, 0.0), c=0.05)
        write_fits_rradtable(tmp_path / "a.fits", t)
        assert rx.lib.rx_rradtable_load_fits(str(tmp_path / "a.fits").encode()) == 0, rx.lib.rx_last_error()
        T = oracle.Tables(rel, lp, rx.rradtable_view())
        _, st_o = oracle.eval_model(oracle.MOD_RELLINELP, T, p, e)
        y, st = rx.eval_model(rx.MODEL_RELLINELP, p, e, precision=rx.FP64, raise_on_error=False)
        print(f"spin above the table: oracle {st_o}, library {st}", flush=True)
        assert st_o != 0 and st != 0 and not np.any(y)
        t = rrad_table(n_spin=3, n_rad=70, n_g=4, c=0.01)
        write_fits_rradtable(tmp_path / "b.fits", t)
        assert rx.lib.rx_rradtable_load_fits(str(tmp_path / "b.fits").encode()) == 0, rx.lib.rx_last_error()
        T = oracle.Tables(rel, lp, rx.rradtable_view())
        # Rout 400: 59 of the 70 zones lie on the disc, within the limit
        ref, st_o = oracle.eval_model(oracle.MOD_RELLINELP, T, p, e)
        y, st = rx.eval_model(rx.MODEL_RELLINELP, p, e, precision=rx.FP64, raise_on_error=False)
        print(f"59 table zones: oracle {st_o}, library {st}, rel {rel_err(y, ref):.3e}", flush=True)
        assert st_o == 0 and st == 0 and rel_err(y, ref) < TOL64
        # a = 0.998 (the table spin) and Rout 1000: all 70
        p3 = list(p)
        p3[2] = 0.998
        p3[5] = 1000.0
        _, st_o = oracle.eval_model(oracle.MOD_RELLINELP, T, p3, e)
        y, st = rx.eval_model(rx.MODEL_RELLINELP, p3, e, precision=rx.FP64, raise_on_error=False)
        print(f"70 table zones: oracle {st_o}, library {st}", flush=True)
        assert st_o == 0 and st != 0 and not np.any(y)
        p2 = list(p)
        p2[9] = 0.0  # without returning radiation the table is not used
        y, st = rx.eval_model(rx.MODEL_RELLINELP, p2, e, precision=rx.FP64)
        assert st == 0 and np.any(y)
    finally:
        rx.rradtable_synthetic(7, 21, 50, 20)
