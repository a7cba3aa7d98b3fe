"""The closed forms of tests/test_oracle_pins.py on the GPU, through the C ABI:
the emission-angle distribution (src/Relprofile.cpp:908-938, :799-801), the
red-wing midpoint rule (:636-644), the lamp-post spin / height interpolation
(src/Rellp.cpp:190-247) and the reflection fractions (src/Rellp.cpp:36-89)."""
import os

import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import (LP_INT, affine_lp_table, dist_expected, dist_table, flat_table, log_grid,
                           lp_affine_value, midpoint_bins, refl_frac_expected, tables, write_lptb, write_rxtb)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tabs():
    rel, lp, T = tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    return T


def _load_rel(tmp_path, rel):
    path = str(tmp_path / "t.rxtb")
    write_rxtb(path, rel)
    assert rx.lib.rx_tables_load(path.encode()) == 0, rx.lib.rx_last_error()


def _load_lp(tmp_path, lp):
    path = str(tmp_path / "lp.rxtb")
    write_lptb(path, lp)
    assert rx.lib.rx_lptable_load(path.encode()) == 0, rx.lib.rx_last_error()


def test_angle_distribution_closed_form(tabs, tmp_path):
    """relxill-style requests on the distribution table with 40 angle bins:
    the device distribution equals the closed form (1e-12) in both
    precisions (the distribution is FP64 in both)."""
    rel, n_incl = dist_table()
    e_ref, refl = np.geomspace(0.07, 1000.0, 301), np.ones((n_incl, 300))
    e_out = log_grid(500, 0.1, 100.0)
    ref = dist_expected(rel["n_g"])
    try:
        _load_rel(tmp_path, rel)
        sid = rx.reflset_create(e_ref, refl)
        reqs = [dict(model=rx.MODEL_RELXILL_CONV, energy=e_out, params=[q, q, 15.0, a, incl, -1.0, 400.0, 0.0],
                     refl_set=sid) for a, incl, q in ((0.5, 40.0, 3.0), (-0.7, 75.0, 1.5), (0.95, 10.0, 5.0))]
        b = rx.Batch(reqs)
        for prec in (rx.FP64, rx.FP32):
            b.run(prec)
            b.sync()
            for i in range(len(reqs)):
                d = b.dist(i, n_incl)
                err = np.max(np.abs(d - ref))
                print(f"dist closed form req {i} prec {prec}: max abs err {err:.3e}", flush=True)
                assert err < 1e-12
                assert np.max(np.abs(d - dist_expected(rel["n_g"], shift=1))) > 1e-3
        b.close()
    finally:
        rx.lib.rx_reflsets_clear()
        rx.tables_synthetic(1, 25, 22, 100, 20)


def test_angle_distribution_per_zone_closed_form(tabs, tmp_path):
    """relxilllp's ten zones on the distribution table and a xillver table of
    40 inclinations: every zone's distribution is the closed form."""
    rel, n_incl = dist_table()
    e = log_grid(400, 0.1, 100.0)
    ref = dist_expected(rel["n_g"])
    try:
        _load_rel(tmp_path, rel)
        rx.xilltable_synthetic(5, 4, 3, 4, 3, n_incl, 300)
        par = [6.0, 0.0, 0.7, 30.0, -1.0, 400.0, 0.0, 2.0, 3.0, 1.0, 300.0, 1.0, 0.0, 0.0]
        b = rx.Batch([dict(model=rx.MODEL_RELXILLLP, energy=e, params=par)])
        b.run(rx.FP64)
        b.sync()
        _, st = b.output(0)
        assert st == 0
        d = b.dist(0, 10 * n_incl).reshape(10, n_incl)
        err = np.max(np.abs(d - ref[None, :]))
        print(f"per-zone dist closed form: max abs err {err:.3e}", flush=True)
        assert err < 1e-12
        b.close()
    finally:
        rx.tables_synthetic(1, 25, 22, 100, 20)
        rx.xilltable_synthetic(5, 4, 3, 4, 3, 10, 600)


@pytest.mark.parametrize("gmin,gmax,n_bins", [(0.5, 0.9, 60), (0.3, 0.94, 40)])
def test_midpoint_rule_below_095(tabs, tmp_path, gmin, gmax, n_bins):
    e = np.linspace(gmin - 0.02, gmax + 0.02, n_bins + 1)
    idx, ref = midpoint_bins(e, gmin, gmax)
    p = [1.0, 3.0, 3.0, 15.0, 0.5, 40.0, -1.0, 400.0, 0.0, 0.0]
    os.environ["RELLINE_PHYSICAL_NORM"] = "1"
    try:
        _load_rel(tmp_path, flat_table(gmin, gmax))
        for prec, tol in ((rx.FP64, 1e-12), (rx.FP32, 1e-5)):
            f, st = rx.eval_model(rx.MODEL_RELLINE, p, e, precision=prec)
            err = np.max(np.abs(f[idx] / ref - 1))
            print(f"midpoint window [{gmin},{gmax}] prec {prec}: max per-bin err {err:.3e}", flush=True)
            assert st == 0 and err < tol
    finally:
        del os.environ["RELLINE_PHYSICAL_NORM"]
        rx.tables_synthetic(1, 25, 22, 100, 20)


LP_POINTS = [(0.5, 7.7, 0.0), (-0.63, 25.3, 0.0), (0.93, 61.0, 0.3), (0.21, 3.4, 0.6), (-0.95, 88.0, 0.1)]


def test_lp_intensity_interpolation_closed_form(tabs, tmp_path):
    """relline_lp with the physical normalisation on the affine lamp-post table
    is A(a, h) (float factors, own height grid per spin) times its value on
    the unit table, bin by bin."""
    e = log_grid(600, 0.1, 10.0)
    lpA = affine_lp_table()
    reqs = [dict(model=rx.MODEL_RELLINELP, energy=e, params=[6.4, h, a, 30.0, -1.0, 400.0, 0.0, 0.0, 2.0, 0.0])
            for a, h, _ in LP_POINTS]
    os.environ["RELLINE_PHYSICAL_NORM"] = "1"
    try:
        out = {}
        for name, lp in (("A", lpA), ("1", affine_lp_table(unit=True))):
            _load_lp(tmp_path, lp)
            b = rx.Batch(reqs)
            for prec in (rx.FP64, rx.FP32):
                b.run(prec)
                b.sync()
                out[name, prec] = [b.output(i) for i in range(len(reqs))]
            b.close()
        for prec, tol in ((rx.FP64, 1e-6), (rx.FP32, 1e-4)):
            for i, (a, h, _) in enumerate(LP_POINTS):
                (fA, sA), (f1, s1) = out["A", prec][i], out["1", prec][i]
                assert sA == 0 and s1 == 0
                sel = f1 > 1e-10 * f1.max()
                A = lp_affine_value(LP_INT, lpA, a, h)
                err = np.max(np.abs(fA[sel] / f1[sel] / A - 1))
                print(f"lp intensity a={a} h={h} prec {prec}: max err {err:.3e}", flush=True)
                assert err < tol
    finally:
        del os.environ["RELLINE_PHYSICAL_NORM"]
        rx.lptable_synthetic(3, 20, 100, 100)


def test_reflection_fractions_closed_form(tabs, tmp_path):
    """relconv_lp requests (moving sources included) on the affine lamp-post
    table: the device's reflection fractions equal the closed form."""
    e = log_grid(300, 0.1, 100.0)
    lp = affine_lp_table()
    flux = np.ones(300)
    reqs = [dict(model=rx.MODEL_RELCONVLP, energy=e, params=[h, beta, a, 30.0, -1.0, 400.0, 0.0, 2.0, 0.0],
                 flux_in=flux) for a, h, beta in LP_POINTS]
    try:
        _load_lp(tmp_path, lp)
        b = rx.Batch(reqs)
        for prec in (rx.FP64, rx.FP32):
            b.run(prec)
            b.sync()
            for i, (a, h, beta) in enumerate(LP_POINTS):
                p, st = oracle.rel_params(oracle.MOD_RELCONVLP, reqs[i]["params"])
                assert st == 0
                ref = refl_frac_expected(lp, a, h, beta, p.rin, p.rout)
                got = b.reflfrac(i)
                err = np.max(np.abs(got / ref - 1))
                print(f"refl frac a={a} h={h} beta={beta} prec {prec}: max rel err {err:.3e}", flush=True)
                assert err < 1e-6
        b.close()
    finally:
        rx.lptable_synthetic(3, 20, 100, 100)
