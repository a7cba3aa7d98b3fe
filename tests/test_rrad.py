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


# ---- relxilllp: the xillver correction factors of the returning radiation ----

@pytest.mark.parametrize("gam", [1.6, 2.0, 2.9])
def test_gshift_boost_closed_forms(gam):
    """corrected_gshift_fluxboost_factor (src/Relreturn_Corona.cpp:39-83): a
    correction c = 1 is the power law's g^Gamma; at g = 1.5 (the shift c is
    defined at) and at g = 2/3 both the parabola (c < 1) and the line (c >= 1)
    give c g^Gamma; below g = 1 the boost is at most 1, and never negative."""
    f = oracle.lib.orc_rrad_gshift_boost
    for g in (0.4, 0.9, 1.0, 1.3, 2.5):
        assert abs(f(1.0, g, gam) / g ** gam - 1) < 1e-14 or (g < 1 and g ** gam > 1)
    for c in (0.35, 0.8, 1.0, 1.25, 1.9):
        assert abs(f(c, 1.5, gam) / (c * 1.5 ** gam) - 1) < 1e-13
        want = min(1.0, c * (2 / 3) ** gam)
        assert abs(f(c, 2 / 3, gam) - want) < 1e-13 * want
    assert f(3.0, 0.9, gam) == 1.0           # clamped: no boost below g = 1
    assert f(0.1, 3.0, gam) == 0.0           # the parabola below zero: clamped


def _power_law_xilltable(gam_t, n_ener=600):
    """A xillver table whose every spectrum (all parameters and inclinations)
    is the bin-integrated photon power law E^-gam_t stored times 10^logXi."""
    from tests.helpers import xill_axes
    ax = xill_axes()
    e = np.geomspace(0.07, 1000.0, n_ener + 1)
    g = (e[1:] ** (1 - gam_t) - e[:-1] ** (1 - gam_t)) / (1 - gam_t)
    L = np.meshgrid(*ax, indexing="ij")[2]
    spec = (10.0 ** L)[..., None] * g
    return dict(n_par=5, pid=[0, 1, 2, 3, 7], n_val=[len(a) for a in ax],
                val=[np.asarray(a, dtype=np.float32) for a in ax], n_ener=n_ener,
                elo=e[:-1].astype(np.float32), ehi=e[1:].astype(np.float32),
                spec=spec.reshape(-1, n_ener).astype(np.float32)), e, g


@pytest.mark.parametrize("gam_t,gam,ect", [(2.0, 2.0, 300.0), (1.7, 2.4, 80.0), (2.6, 1.9, 600.0)])
def test_correction_factors_of_a_power_law_table(gam_t, gam, ect):
    """get_xillver_fluxcorrection_factors (src/Xillspec.cpp:500-526) on a table
    of identical photon power laws E^-gam_t: (1) the spectrum shifted to 1.5
    times lower energies carries 1.5^-gam_t of the photons of any band it
    covers, so the g-shift factor is 1.5^(gam_t - Gamma) (Gamma the model's
    index); a shift the other way, or no time-dilation factor 2/3, misses;
    (2) the flux factor is W times the band energy flux of the spectrum over
    that of the normalised primary, which is 1e15 / (4 pi 1e20 1.602177e-9)
    (the normalisation's definition, src/Xillspec.cpp:179-205); W is the sum
    of the angle weights 0.5 cos(i)/n_incl of calc_angle_averaged_xill_spec,
    where the table's inclination (degrees) is read as radians."""
    x, e, g = _power_law_xilltable(gam_t)
    xt = oracle.Xill(x)
    xp = oracle.xill_param(gam, 3.0, 1.0, ect)
    cf, cg = oracle.xill_corr_factors(xt, xp)
    want_g = 1.5 ** (gam_t - gam)
    assert abs(cg / want_g - 1) < 2e-3
    assert abs(cg / 1.5 ** (-gam_t - gam) - 1) > 0.5 and abs(cg / (want_g * 1.5) - 1) > 0.2
    incl = np.asarray(x["val"][4], dtype=np.float64)
    n_incl = incl.size
    W = np.sum(0.5 * np.cos(incl) / n_incl)
    W_deg = np.sum(0.5 * np.cos(np.radians(incl)) / n_incl)
    ef = e.astype(np.float32).astype(np.float64)
    em = 0.5 * (ef[1:] + ef[:-1])
    gf = g.astype(np.float32).astype(np.float64)
    band = (ef[:-1] >= 0.1) & (ef[1:] <= 1000.0)
    q = 1e15 / (4 * np.pi * 1e20 * 1.602177e-9)
    want_f = W * np.sum(gf[band] * em[band]) / q
    assert abs(cf / want_f - 1) < 1e-3
    assert abs(W_deg / W - 1) > 0.1  # the degree reading would be visible


def test_returning_emissivity_with_correction_factors():
    """The correction factors in calc_rrad_emis_corona (src/Relreturn_Corona.cpp
    :119-151): with every returning photon at g = 1.5, a g-shift factor c turns
    the weight 1.5^(Gamma-1) into boost(c, 1.5)/1.5 = c 1.5^(Gamma-1), and the
    flux factor multiplies the zone's sum: equal factors in every zone scale
    the returning emissivity by c times the flux factor at every radius;
    c within 1e-3 of 1 leaves the plain weight."""
    tf = [[0.03, 0.07], [0.011, 0.005]]
    t = two_zone_table(tf, g0=1.5)
    p, st = oracle.rel_params(oracle.MOD_RELLINELP, LP_P)
    re = oracle.fine_radial_grid(p.rin, p.rout, 1000)
    e0 = re ** -3.0
    er0, st = oracle.rrad_emis(t, p, re, e0)
    assert st == 0
    rg = np.array([p.rin, 30.0, 200.0, p.rout])
    for c, fl in ((0.7, 2.5), (1.4, 0.8), (1.0005, 1.7)):
        cf = np.full(3, fl)
        cg = np.full(3, c)
        p.rr_corr_flux = cf.ctypes.data_as(oracle._D)
        p.rr_corr_gshift = cg.ctypes.data_as(oracle._D)
        p.rr_corr_rgrid = rg.ctypes.data_as(oracle._D)
        p.rr_corr_n = 3
        er, st = oracle.rrad_emis(t, p, re, e0)
        assert st == 0
        want = fl * (c if abs(c - 1) > 1e-3 else 1.0)
        assert np.abs(er / er0 / want - 1).max() < 1e-12, c
    p.rr_corr_n = 0


def test_correction_factors_follow_the_zone_of_each_table_centre():
    """rebin_corrfactors_to_rradtable_grid (src/Relreturn_Corona.cpp:248-289):
    a table zone takes the factors of the relxill zone its centre falls in.
    Two table zones whose centres fall in different relxill zones: the
    returning emissivity at each centre (the log-linear profile's value
    there) is scaled by that zone's flux factor; one zone off would swap
    them."""
    tf = [[0.03, 0.07], [0.011, 0.005]]
    t = two_zone_table(tf, g0=1.0)
    p, st = oracle.rel_params(oracle.MOD_RELLINELP, LP_P)
    re = oracle.fine_radial_grid(p.rin, p.rout, 1000)
    e0 = re ** -3.0
    er0, st = oracle.rrad_emis(t, p, re, e0)
    assert st == 0
    rad = np.array([0.5 * (re[-1] + 20.0), 0.5 * (20.0 + re[0])])
    rg = np.array([p.rin, 10.0, 100.0, 400.0, p.rout])  # centres (12.1, 510) in zones 1 and 3
    cf = np.array([5.0, 2.0, 7.0, 3.0])
    cg = np.ones(4)
    p.rr_corr_flux = cf.ctypes.data_as(oracle._D)
    p.rr_corr_gshift = cg.ctypes.data_as(oracle._D)
    p.rr_corr_rgrid = rg.ctypes.data_as(oracle._D)
    p.rr_corr_n = 4
    er, st = oracle.rrad_emis(t, p, re, e0)
    p.rr_corr_n = 0
    assert st == 0
    c0a, c1a = _line_through(re, er)
    c0b, c1b = _line_through(re, er0)
    ratio = np.exp((c0a - c0b) + (c1a - c1b) * rad)
    want = np.array([cf[np.searchsorted(rg, x) - 1] for x in rad])
    assert list(want) == [2.0, 3.0]
    assert np.abs(ratio / want - 1).max() < 1e-9
