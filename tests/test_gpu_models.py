"""GPU tests beyond single-model parity: the relxill-style angle-resolved
convolution against the oracle, the closed form of the r/g* integration on a
flat transfer function, continuous-batching isolation (a request's result does
not depend on what it is batched with), and the FP64 re-check counter path."""
import os

import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import (RELLINE_CFG0, abs_err, draw_lp_params, draw_relconv_params, flat_table, log_grid,
                           reflection_spectrum, rel_err, tables, write_rxtb)

pytestmark = pytest.mark.gpu

TOL64 = 1e-8
TOL32 = 1e-4


@pytest.fixture(scope="module")
def tabs():
    rel, lp, T = tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    return T


def test_relxill_style_angle_resolved_conv(tabs):
    """relxill's per-angle convolution (src/Relxill.cpp:289-482 with the
    angle distribution of src/Relprofile.cpp:908-938): 10 cos-uniform angle
    bins, reflection spectra on 2999 bins over 0.07-1000 keV, profile grid of
    4096 log bins over 0.01-1000 keV, output on 3000 bins over 0.1-100 keV."""
    cg = log_grid(4096, 0.01, 1000.0)
    rx.set_conv_grid(4096, 0.01, 1000.0)
    try:
        e_ref, refl = reflection_spectrum(n_incl=10)
        sid = rx.reflset_create(e_ref, refl)
        e_out = log_grid(3000, 0.1, 100.0)
        for p in ([3.0, 3.0, 15.0, 0.998, 30.0, -1.0, 400.0, 0.0],
                  [2.0, 4.0, 8.0, -0.5, 70.0, 3.0, 300.0, 0.0]):
            ref, dist_ref, st = oracle.relxill_conv(tabs, p, cg, e_ref, refl, e_out)
            assert st == 0
            b = rx.Batch([dict(model=rx.MODEL_RELXILL_CONV, energy=e_out, params=p, refl_set=sid)])
            for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
                b.run(prec)
                b.sync()
                y, st = b.output(0)
                assert st == 0
                d = b.dist(0, 10)
                print(f"relxill-style {p[3:5]} prec {prec}: rel {rel_err(y, ref):.3e} dist {rel_err(d, dist_ref):.3e}",
                      flush=True)
                assert rel_err(y, ref) < tol
                assert rel_err(d, dist_ref) < tol
            b.close()
    finally:
        rx.lib.rx_reflsets_clear()
        rx.set_conv_grid(4096, 0.00035, 2000.0)


def test_flat_transfer_function_closed_form(tabs, tmp_path):
    """With trff_1 + trff_2 = 1 on a 2e-3 window around g = 1 and the physical
    normalisation, the photon flux is pi/4 (int_0^1 dg*/sqrt(g*(1-g*)) times
    the normalised emissivity); both precisions, against the same closed form
    the oracle meets."""
    e = np.linspace(0.998, 1.002, 4001)
    p = [1.0, 3.0, 3.0, 15.0, 0.5, 40.0, -1.0, 400.0, 0.0, 0.0]

    def run():
        for prec in (rx.FP64, rx.FP32):
            f, st = rx.eval_model(rx.MODEL_RELLINE, p, e, precision=prec)
            print(f"flat table prec {prec}: sum {f.sum():.9f} pi/4 {np.pi / 4:.9f}", flush=True)
            assert abs(f.sum() / (np.pi / 4) - 1.0) < 2e-3
    _with_table(tmp_path, flat_table(0.999, 1.001), run)


def _mixed_requests(n, seed):
    rng = np.random.default_rng(seed)
    e_ref, refl = reflection_spectrum(n_incl=1)
    reqs = []
    lp = draw_lp_params(n, seed=seed + 1)
    cv = draw_relconv_params(n, seed=seed + 2)
    for k in range(n):
        nb = int(np.exp(rng.uniform(np.log(300), np.log(4000))))
        u = k % 3
        if u == 0:
            p = list(RELLINE_CFG0)
            p[4], p[5] = cv[k][3], cv[k][4]
            reqs.append(dict(model=rx.MODEL_RELLINE, energy=log_grid(nb, 0.1, 10.0), params=p))
        elif u == 1:
            e = log_grid(nb, 0.1, 100.0)
            reqs.append(dict(model=rx.MODEL_RELCONV, energy=e, params=cv[k], flux_in=oracle.rebin(e, e_ref, refl[0])))
        else:
            reqs.append(dict(model=rx.MODEL_RELLINELP, energy=log_grid(nb, 0.1, 10.0), params=lp[k]))
    return reqs


def _same_up_to_sum_order(y, solo, model):
    """A request's output against its solo run: only the order of the FP64
    atomic adds of the per-chunk sums differs (~1e-16 of each bin).  Line
    models keep that per bin; the convolution's FFT spreads it over the
    output, so there the bound is ~1e-16 of the spectrum's maximum."""
    if model in (rx.MODEL_RELLINE, rx.MODEL_RELLINELP):
        return rel_err(y, solo) < 1e-12
    return abs_err(y, solo) < 1e-13


def test_batch_equals_solo_runs(tabs):
    """A request's output in a mixed batch equals its solo evaluation: its work
    items depend only on itself; only the FP64 order of the per-chunk sums may
    differ."""
    reqs = _mixed_requests(24, 31)
    b = rx.Batch(reqs)
    for prec in (rx.FP64, rx.FP32):
        b.run(prec)
        b.sync()
        for i, r in enumerate(reqs):
            y, st = b.output(i)
            assert st == 0
            solo, st = rx.eval_model(r["model"], r["params"], r["energy"], flux_in=r.get("flux_in"), precision=prec)
            assert _same_up_to_sum_order(y, solo, r["model"]), (prec, i, rel_err(y, solo), abs_err(y, solo))
    b.close()


def test_work_item_shape_does_not_change_results(tabs):
    """Bins are cut into work items of S bins; a Romberg bin reuses the
    integrand pair its predecessor left at the shared edge, except at an
    item's first bin, which evaluates it afresh.  Both give the same bits, so
    line-model outputs agree per bin up to the FP64 order of the atomic adds
    whatever S is."""
    reqs = [r for r in _mixed_requests(30, 57) if r["model"] in (rx.MODEL_RELLINE, rx.MODEL_RELLINELP)]
    b = rx.Batch(reqs)
    try:
        for prec in (rx.FP64, rx.FP32):
            outs = []
            for seg, wpc in ((32, 8), (256, 4), (128, 16)):
                b.set_shape(seg, wpc)
                b.run(prec)
                b.sync()
                outs.append([b.output(i)[0].copy() for i in range(len(reqs))])
            worst = max(rel_err(o[i], outs[0][i]) for o in outs[1:] for i in range(len(reqs)))
            print(f"work item shapes prec {prec}: worst per-bin difference {worst:.3e}", flush=True)
            assert worst < 1e-12
    finally:
        b.close()


def test_server_isolation_and_latency(tabs):
    """Continuous batching: requests submitted while others run come back
    equal to their solo runs (FP32) with per-request latency recorded."""
    reqs = _mixed_requests(60, 41)
    srv = rx.Server(max_inflight=16, hbm_budget=1 << 30, n_streams=3, precision=rx.FP32)
    try:
        tickets = [srv.submit(r["model"], r["energy"], r["params"], flux_in=r.get("flux_in")) for r in reqs]
        srv.drain()
        st = srv.stats()
        assert st["completed"] == len(reqs)
        assert st["batches"] >= 4  # the in-flight cap split the stream into several batches
        for t, r in zip(tickets, reqs):
            y, status, lat = srv.poll(t)
            assert status == 0 and lat > 0
            solo, _ = rx.eval_model(r["model"], r["params"], r["energy"], flux_in=r.get("flux_in"), precision=rx.FP32)
            assert _same_up_to_sum_order(y, solo, r["model"]), (rel_err(y, solo), abs_err(y, solo))
    finally:
        srv.stop()


def test_server_reports_bad_requests_without_disturbing_others(tabs):
    srv = rx.Server(max_inflight=64, hbm_budget=1 << 30, n_streams=2, precision=rx.FP32)
    try:
        e = log_grid(500, 0.1, 10.0)
        good = srv.submit(rx.MODEL_RELLINE, e, RELLINE_CFG0)
        bad_p = list(RELLINE_CFG0)
        bad_p[6], bad_p[7] = 50.0, 40.0  # Rout <= Rin
        bad = srv.submit(rx.MODEL_RELLINE, e, bad_p)
        srv.drain()
        y, st, _ = srv.poll(good)
        yb, stb, _ = srv.poll(bad)
        assert st == 0 and stb == 1  # RX_ERR_PARAM
        assert np.all(yb == 0)
        solo, _ = rx.eval_model(rx.MODEL_RELLINE, RELLINE_CFG0, e, precision=rx.FP32)
        assert rel_err(y, solo) < 1e-12
    finally:
        srv.stop()


def _with_table(tmp_path, rel, fn):
    path = str(tmp_path / "t.rxtb")
    write_rxtb(path, rel)
    os.environ["RELLINE_PHYSICAL_NORM"] = "1"
    try:
        assert rx.lib.rx_tables_load(path.encode()) == 0, rx.lib.rx_last_error()
        return fn()
    finally:
        del os.environ["RELLINE_PHYSICAL_NORM"]
        rx.tables_synthetic(1, 25, 22, 100, 20)


def test_sawtooth_per_bin_closed_form(tabs, tmp_path):
    """The swapped-lerp sawtooth, bin by bin against quadrature, on the GPU."""
    from tests.helpers import sawtooth_bins, sawtooth_table
    rel = sawtooth_table()
    e = np.linspace(0.998, 1.002, 4001)
    p = [1.0, 3.0, 3.0, 15.0, 0.5, 40.0, -1.0, 400.0, 0.0, 0.0]
    idx, ref = sawtooth_bins(e, rel)

    def run():
        for prec, tol in ((rx.FP64, 1e-6), (rx.FP32, 1e-4)):
            f, st = rx.eval_model(rx.MODEL_RELLINE, p, e, precision=prec)
            err = np.max(np.abs(f[idx] / ref - 1))
            print(f"sawtooth prec {prec}: max per-bin error {err:.3e}", flush=True)
            assert err < tol
    _with_table(tmp_path, rel, run)


@pytest.mark.parametrize("a,incl", [(0.5, 40.0), (-0.63, 71.0)])
def test_affine_table_closed_form(tabs, tmp_path, a, incl):
    """Bilinear (a, mu0) interpolation with float factors, on the GPU."""
    from tests.helpers import affine_expected, affine_table, sawtooth_bins
    rel = affine_table()
    e = np.linspace(0.998, 1.002, 4001)
    k, _ = affine_expected(rel, a, incl)
    idx, unit = sawtooth_bins(e, rel, c_rel=0.0)
    p = [1.0, 3.0, 3.0, 15.0, a, incl, -1.0, 400.0, 0.0, 0.0]

    def run():
        for prec, tol in ((rx.FP64, 1e-6), (rx.FP32, 1e-4)):
            f, st = rx.eval_model(rx.MODEL_RELLINE, p, e, precision=prec)
            assert np.max(np.abs(f[idx] / (k * unit) - 1)) < tol
    _with_table(tmp_path, rel, run)


def test_radial_mapping_closed_form(tabs, tmp_path):
    """Node radii and fine-grid lerp on the GPU: flux ratio of the radial
    table to the flat table equals the weighted mean of (u + v re/1000)."""
    from tests.helpers import radial_expected, radial_table
    e = np.linspace(0.998, 1.002, 4001)
    p = [1.0, 3.0, 3.0, 15.0, 0.5, 40.0, -2.0, 400.0, 0.0, 0.0]
    expect = radial_expected(2 * oracle.lib.orc_kerr_rms(0.5), 400.0)
    for prec, tol in ((rx.FP64, 1e-7), (rx.FP32, 1e-4)):
        fr = _with_table(tmp_path, radial_table(), lambda: rx.eval_model(rx.MODEL_RELLINE, p, e, precision=prec)[0])
        ff = _with_table(tmp_path, flat_table(0.999, 1.001),
                         lambda: rx.eval_model(rx.MODEL_RELLINE, p, e, precision=prec)[0])
        got = fr.sum() / ff.sum()
        print(f"radial mapping prec {prec}: {got:.10f} expected {expect:.10f}", flush=True)
        assert abs(got / expect - 1) < tol


def test_isco_snap_table_parity(tabs, tmp_path):
    """The ISCO snap (src/Relprofile.cpp:195-200) on the GPU: a table whose
    innermost radius sits 5e-4 above the ISCO, at a table spin, bin by bin
    against the oracle."""
    from tests.helpers import isco_snap_table, rel_err
    rel = isco_snap_table()
    a = float(rel["a"][int(np.searchsorted(rel["a"], 0.5))])
    e = np.linspace(0.998, 1.002, 4001)
    p = [1.0, 3.0, 3.0, 15.0, a, 40.0, -1.0, 20.0, 0.0, 0.0]
    ref, st = oracle.eval_model(oracle.MOD_RELLINE, oracle.Tables(rel), p, e, phys_norm=True)
    assert st == 0

    def run():
        for prec, tol in ((rx.FP64, 1e-8), (rx.FP32, 1e-4)):
            f, st = rx.eval_model(rx.MODEL_RELLINE, p, e, precision=prec)
            err = rel_err(f, ref)
            print(f"isco snap table prec {prec}: status {st} max rel err {err:.3e}", flush=True)
            assert st == 0 and err < tol
    _with_table(tmp_path, rel, run)


def test_lp_switch_table_parity(tabs, tmp_path):
    """The 75 deg lin/log radius factor of the lamp-post emissivity
    (src/relutility.c:407-417, src/Rellp.cpp:120-175) on the GPU: a lamp-post table whose
    emission angle crosses 75 deg at 10 rg, relline_lp bin by bin against the
    oracle."""
    from tests.helpers import rel_err, switch_lp_table, write_lptb
    rel, lp = switch_lp_table(10.0)
    a = float(lp["a"][11])
    e = log_grid(3000, 0.1, 10.0)
    p = [6.4, 6.0, a, 30.0, -1.0, 100.0, 0.0, 0.0, 2.0, 0.0]
    ref, st = oracle.eval_model(oracle.MOD_RELLINELP, oracle.Tables(rel, lp), p, e)
    assert st == 0
    path = str(tmp_path / "lp.rxtb")
    write_lptb(path, lp)
    try:
        assert rx.lib.rx_lptable_load(path.encode()) == 0, rx.lib.rx_last_error()
        for prec, tol in ((rx.FP64, 1e-8), (rx.FP32, 1e-4)):
            f, st = rx.eval_model(rx.MODEL_RELLINELP, p, e, precision=prec)
            err = rel_err(f, ref)
            print(f"lp switch table prec {prec}: status {st} max rel err {err:.3e}", flush=True)
            assert st == 0 and err < tol
    finally:
        rx.lptable_synthetic(3, 20, 100, 100)
