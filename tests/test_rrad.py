"""Returning radiation of the lamp-post emissivity on the CPU: the oracle's
restatement of get_rrad_emis_corona (src/Relreturn_Corona.cpp:291-321) with
the table selection and trimming of src/Relreturn_Table.cpp:343-480, pinned by
closed forms on tables with two zones, where the result is exactly log-linear
in radius; the library's table loader.  No GPU needed."""
import math

import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import rrad_table, write_fits_rradtable

LP_P = [6.4, 6.0, 0.5, 30.0, -1.0, 1000.0, 0.0, 0.0, 2.0, 1.0]  # relline_lp with switch_returnrad 1


def two_zone_table(tf, spins=(-0.5, 0.5, 0.998), g0=1.0, r_mid=20.0, n_g=4):
    """Two zones per spin: [ISCO, r_mid] and [r_mid, 1000]; tf[ro][re] the
    returning fractions, all of it at the energy shift g0."""
    t = rrad_table(n_spin=len(spins), n_rad=2, n_g=n_g, spins=spins, c=0.0, g0=g0)
    ns = len(spins)
    rlo = np.asarray(t["rlo"]).reshape(ns, 2)
    rhi = np.asarray(t["rhi"]).reshape(ns, 2)
    rlo[:, 1] = r_mid
    rhi[:, 0] = r_mid
    tt = np.zeros((ns, 2, 2))
    for k in range(ns):
        tt[k] = np.asarray(tf[k] if np.ndim(tf) == 3 else tf)
    t.update(rlo=rlo.ravel(), rhi=rhi.ravel(), tf_r=tt.ravel())
    return t


def _line_through(re, y):
    """log y is linear in r on the whole grid for two zones: its coefficients"""
    ly = np.log(y)
    c1, c0 = np.polyfit(re, ly, 1)
    assert np.abs(ly - (c0 + c1 * re)).max() < 1e-12 * np.abs(ly).max() + 1e-12
    return c0, c1


def _run(t, params, emis0_fn):
    p, st = oracle.rel_params(oracle.MOD_RELLINELP, params)
    assert st == 0
    re = oracle.fine_radial_grid(p.rin, p.rout, 1000)
    er, st = oracle.rrad_emis(t, p, re, emis0_fn(re))
    assert st == 0
    return re, er


def test_two_zones_return_the_table_fractions():
    """With a linear direct emissivity e0(r) = 1 + r/1000 (which inv_rebin_mean
    interpolates exactly), returning fractions tf[ro][re] and g = 1: the zone
    centres receive sum_re tf[ro][re] e0(centre re), and the fine-grid profile is
    the log-linear line through them.  Swapped ro / re or a missed end
    replacement of the zone edges fail."""
    tf = [[0.03, 0.07], [0.011, 0.005]]
    t = two_zone_table(tf)
    e0 = lambda r: 1.0 + r / 1000.0
    re, er = _run(t, LP_P, e0)
    c0, c1 = _line_through(re, er)
    rad = np.array([0.5 * (re[-1] + 20.0), 0.5 * (20.0 + re[0])])
    ret = np.exp(c0 + c1 * rad)
    want = np.array([tf[0][0] * e0(rad[0]) + tf[0][1] * e0(rad[1]), tf[1][0] * e0(rad[0]) + tf[1][1] * e0(rad[1])])
    assert np.abs(ret / want - 1).max() < 1e-9


def test_energy_shift_weight_is_g_to_gamma_minus_one():
    """All returning photons at g = 1.3: the profile is 1.3^(Gamma-1) times the
    one at g = 1 (calc_rrad_emis_zone); g within 1e-3 of 1 is not boosted."""
    tf = [[0.03, 0.07], [0.011, 0.005]]
    e0 = lambda r: r ** -3.0
    _, er1 = _run(two_zone_table(tf, g0=1.0), LP_P, e0)
    for gam in (2.0, 2.7):
        p = list(LP_P)
        p[8] = gam
        _, er13 = _run(two_zone_table(tf, g0=1.3), p, e0)
        _, er1g = _run(two_zone_table(tf, g0=1.0), p, e0)
        assert np.abs(er13 / er1g / 1.3 ** (gam - 1) - 1).max() < 1e-12
        _, er_near = _run(two_zone_table(tf, g0=1.0 + 5e-4), p, e0)
        assert np.abs(er_near / er1g - 1).max() < 1e-12


def test_spin_selection_and_inner_area_correction():
    """The first tabulated spin at or above a is used (spins -0.5, 0.5, 0.998
    with fractions scaled 1, 2, 3: a = 0.3 takes 0.5's); the inner zone's
    fractions from it are scaled by the area of [Rin, r_mid] over that of the
    table's [ISCO(0.5), r_mid]."""
    base = np.array([[0.03, 0.07], [0.011, 0.005]])
    t = two_zone_table([base, 2 * base, 3 * base])
    e0 = lambda r: 1.0 + r / 1000.0
    p = list(LP_P)
    p[2] = 0.3
    re, er = _run(t, p, e0)
    c0, c1 = _line_through(re, er)
    rin, rmid = re[-1], 20.0
    rad = np.array([0.5 * (rin + rmid), 0.5 * (rmid + re[0])])
    ret = np.exp(c0 + c1 * rad)
    r_tab = oracle.lib.orc_kerr_rms(0.5)
    corr = (0.5 * (rin + rmid) * (rmid - rin)) / (0.5 * (r_tab + rmid) * (rmid - r_tab))
    tf = 2 * base
    want = np.array([tf[0][0] * corr * e0(rad[0]) + tf[0][1] * e0(rad[1]), tf[1][0] * e0(rad[0]) + tf[1][1] * e0(rad[1])])
    assert np.abs(ret / want - 1).max() < 1e-9


def test_relline_lp_with_returning_radiation_adds_it():
    """switch_returnrad = 1 adds the returning emissivity to the direct one
    (add_returnrad_emis): with physical normalisation off the line is
    renormalised, so the added part shows as a change of shape only where
    the profile differs; relline_lp's profile with it equals the profile of
    the summed emissivity (checked through the system parameters)."""
    from tests.helpers import tables
    rel, lp, _ = tables()
    rx.rradtable_synthetic(7, 21, 50, 20)
    rr = rx.rradtable_view()
    T = oracle.Tables(rel, lp, rr)
    p0, _ = oracle.rel_params(oracle.MOD_RELLINELP, [6.4, 6.0, 0.9, 30, -1, 400, 0, 0, 2, 0])
    p1, _ = oracle.rel_params(oracle.MOD_RELLINELP, [6.4, 6.0, 0.9, 30, -1, 400, 0, 0, 2, 1])
    s0, st0 = oracle.system_parameters(T, p0)
    s1, st1 = oracle.system_parameters(T, p1)
    assert st0 == 0 and st1 == 0
    er, st = oracle.rrad_emis(rr, p0, s0["re"], s0["emis"])
    assert st == 0 and er.min() > 0
    assert np.array_equal(s1["emis"], s0["emis"] + er)
    p2, _ = oracle.rel_params(oracle.MOD_RELLINELP, [6.4, 6.0, 0.9, 30, -1, 400, 0, 0, 2, 2])
    s2, _ = oracle.system_parameters(T, p2)
    assert np.array_equal(s2["emis"], er)


def test_library_rradtable_loader_round_trip(tmp_path):
    t = rrad_table(n_spin=3, n_rad=7, n_g=5, c=0.04, shift=1)
    write_fits_rradtable(tmp_path / "table_returnRad_v20220301.fits", t)
    try:
        assert rx.lib.rx_rradtable_load_fits(str(tmp_path / "table_returnRad_v20220301.fits").encode()) == 0, \
            rx.lib.rx_last_error()
        v = rx.rradtable_view()
        assert (v["n_spin"], v["n_rad"], v["n_g"]) == (3, 7, 5)
        for k in ("spin", "rlo", "rhi", "tf_r", "gmin", "gmax", "frac_g"):
            assert np.array_equal(v[k], np.asarray(t[k], dtype=np.float64)), k
    finally:
        rx.rradtable_synthetic()
