"""Closed-form pins for four oracle functions that the known-answer tests and
the flat / sawtooth / affine tables leave open: the emission-angle
distribution of calc_relline_profile (src/Relprofile.cpp:908-938, with
get_cosne_bin :799-801), the red-wing midpoint rule of int_romb (:636-644),
the lamp-post spin / height interpolation (interpol_lptable,
src/Rellp.cpp:190-247, get_ipol_factor src/relutility.c:419-423) and the
reflection fractions (calc_refl_frac, src/Rellp.cpp:36-89).  Each check is
chosen so that a plausible mistake (a bin off by one, a dropped weight, a
Romberg or lower-end rule, a height factor from the wrong spin, exchanged
spin weights, a missed or reversed aberration) fails it.  CPU only."""
import numpy as np
import pytest

import oracle
from tests.helpers import (LP_DEL, LP_DEL_S, LP_INC, LP_INC_S, LP_INT, affine_lp_table, dist_expected, dist_table,
                           flat_table, log_grid, lp_affine_value, midpoint_bins, refl_frac_expected, tables)


# ---- angle distribution ----------------------------------------------------

@pytest.mark.parametrize("a,incl,q", [(0.5, 40.0, 3.0), (-0.7, 75.0, 1.5), (0.95, 10.0, 5.0)])
def test_angle_distribution_closed_form(a, incl, q):
    rel, n_incl = dist_table()
    T = oracle.Tables(rel)
    p, st = oracle.rel_params(oracle.MOD_RELCONV, [q, q, 15.0, a, incl, -1.0, 400.0, 0.0])
    assert st == 0
    cg = oracle.log_grid(4097, 0.00035, 2000.0)
    _, dist, st = oracle.relbase(T, p, cg, n_incl=n_incl)
    assert st == 0
    ref = dist_expected(rel["n_g"])
    assert np.max(np.abs(dist - ref)) < 1e-12
    # the mistakes: a bin off by one, d_gstar or g^2 dropped, branches exchanged
    for bad in (dist_expected(rel["n_g"], shift=1), dist_expected(rel["n_g"], drop_dgstar=True),
                dist_expected(rel["n_g"], drop_g2=True), dist_expected(rel["n_g"], c1=0.3, c2=0.7)):
        assert np.max(np.abs(dist - bad)) > 1e-3


def test_angle_distribution_per_zone_closed_form():
    """Ten lamp-post zones (relxilllp's radial grid): every zone's own
    normalised distribution is the same closed form, the radial factor of
    each zone cancelling in its normalisation."""
    rel, n_incl = dist_table()
    _, lp, _ = tables()
    T = oracle.Tables(rel, lp)
    p, st = oracle.rel_params(oracle.MOD_RELCONVLP, [6.0, 0.0, 0.7, 30.0, -1.0, 400.0, 0.0, 2.0, 0.0])
    assert st == 0
    p.num_zones = 10
    cg = oracle.log_grid(4097, 0.00035, 2000.0)
    flux = np.zeros(4096 * 10)
    dist = np.zeros(10 * n_incl)
    import ctypes
    st = oracle.lib.orc_relbase(ctypes.byref(T.rel), T.lp_ptr, ctypes.byref(p), cg.ctypes.data_as(oracle._D),
                                4096, 1000, flux.ctypes.data_as(oracle._D), dist.ctypes.data_as(oracle._D),
                                n_incl, None)
    assert st == 0
    ref = dist_expected(rel["n_g"])
    for z in range(10):
        assert np.max(np.abs(dist[z * n_incl:(z + 1) * n_incl] - ref)) < 1e-12, z


# ---- red-wing midpoint rule ------------------------------------------------

@pytest.mark.parametrize("gmin,gmax,n_bins", [(0.5, 0.9, 60), (0.3, 0.94, 40)])
def test_midpoint_rule_below_095(gmin, gmax, n_bins):
    """A flat window entirely below 0.95 line energies: every bin's interior
    integral is the midpoint rule, so each interior bin equals f(mid) (hi -
    lo) to rounding; the quadrature (Romberg everywhere) and the lower-end
    rule are measurably different."""
    rel = flat_table(gmin, gmax)
    T = oracle.Tables(rel)
    e = np.linspace(gmin - 0.02, gmax + 0.02, n_bins + 1)
    f, st, s = oracle.eval_model(oracle.MOD_RELLINE, T, [1.0, 3.0, 3.0, 15.0, 0.5, 40.0, -1.0, 400.0, 0.0, 0.0],
                                 e, phys_norm=True, stats=True)
    assert st == 0
    assert s["romberg_calls"] == 0 and s["midpoint_calls"] > 0
    idx, ref = midpoint_bins(e, gmin, gmax)
    assert idx.size > n_bins // 2
    assert np.max(np.abs(f[idx] / ref - 1)) < 1e-12
    _, quad = midpoint_bins(e, gmin, gmax, rule="quad")
    _, low = midpoint_bins(e, gmin, gmax, rule="lo")
    assert np.max(np.abs(quad / ref - 1)) > 1e-5
    assert np.max(np.abs(low / ref - 1)) > 1e-2


def test_red_wing_switch_at_095():
    """A window straddling 0.95: bins starting at or above 0.95 take Romberg,
    the others the midpoint rule; the midpoint closed form holds exactly for
    the bins below, and those above stay within Romberg's 2% of the
    quadrature but off the midpoint value."""
    gmin, gmax = 0.6, 1.3
    rel = flat_table(gmin, gmax)
    T = oracle.Tables(rel)
    e = np.linspace(0.58, 1.32, 75)
    f, st = oracle.eval_model(oracle.MOD_RELLINE, T, [1.0, 3.0, 3.0, 15.0, 0.5, 40.0, -1.0, 400.0, 0.0, 0.0],
                              e, phys_norm=True)
    assert st == 0
    idx, mid = midpoint_bins(e, gmin, gmax)
    _, quad = midpoint_bins(e, gmin, gmax, rule="quad")
    below = e[idx] < 0.95
    above = e[idx] >= 0.95
    assert below.sum() > 10 and above.sum() > 10
    assert np.max(np.abs(f[idx][below] / mid[below] - 1)) < 1e-12
    assert np.max(np.abs(f[idx][above] / quad[above] - 1)) < 2e-2
    assert np.max(np.abs(f[idx][above] / mid[above] - 1)) > 1e-6


# ---- lamp-post spin / height interpolation and reflection fractions --------

LP_POINTS = [(0.5, 7.7, 0.0), (-0.63, 25.3, 0.0), (0.93, 61.0, 0.3), (0.21, 3.4, 0.6), (-0.95, 88.0, 0.1)]


def _lp_sys(lp, a, h, beta, rin=-1.0, rout=400.0):
    rel, _, _ = tables()
    T = oracle.Tables(rel, lp)
    p, st = oracle.rel_params(oracle.MOD_RELCONVLP, [h, beta, a, 30.0, rin, rout, 0.0, 2.0, 0.0])
    assert st == 0
    sp, st = oracle.system_parameters(T, p)
    assert st == 0
    return sp, p


@pytest.mark.parametrize("a,h,beta", LP_POINTS)
def test_lp_spin_height_interpolation_closed_form(a, h, beta):
    """Intensity, emission and incidence angle of the lamp-post table
    interpolated in spin and height: the intensity ratio to the unit table is
    A(a, h) at every fine radius, the angles are D(a, h) + s ln r, each with
    the float factors of the reference; a height factor taken from the lower
    spin for both, or exchanged spin weights, miss by > 1e-3."""
    lpA, lp1 = affine_lp_table(), affine_lp_table(unit=True)
    spA, p = _lp_sys(lpA, a, h, beta)
    sp1, _ = _lp_sys(lp1, a, h, beta)
    A = lp_affine_value(LP_INT, lpA, a, h)
    D = lp_affine_value(LP_DEL, lpA, a, h)
    Di = lp_affine_value(LP_INC, lpA, a, h)
    re = spA["re"]
    assert np.max(np.abs(spA["emis"] / sp1["emis"] / A - 1)) < 1e-6
    assert np.max(np.abs(spA["del_emit"] - (D + LP_DEL_S * np.log(re)))) < 2e-6
    assert np.max(np.abs(spA["del_inc"] - (Di + LP_INC_S * np.log(re)))) < 2e-6
    for mistake in ("one_h", "swap_a"):
        assert abs(lp_affine_value(LP_INT, lpA, a, h, mistake) / A - 1) > 1e-3, mistake
        assert abs(lp_affine_value(LP_DEL, lpA, a, h, mistake) - D) > 1e-4, mistake


@pytest.mark.parametrize("a,h,beta", LP_POINTS)
def test_reflection_fractions_closed_form(a, h, beta):
    lp = affine_lp_table()
    sp, p = _lp_sys(lp, a, h, beta)
    ref = refl_frac_expected(lp, a, h, beta, p.rin, p.rout)
    assert np.max(np.abs(sp["refl"] / ref - 1)) < 1e-6, (sp["refl"], ref)
    if beta > 0:
        # the aberration left out, or taken with the opposite sign
        plain = refl_frac_expected(lp, a, h, 0.0, p.rin, p.rout)
        assert abs(plain[0] / ref[0] - 1) > 1e-2
        rev = refl_frac_expected(lp, a, h, -beta, p.rin, p.rout)
        assert abs(rev[0] / ref[0] - 1) > 1e-2


def test_reflection_fraction_uses_the_largest_table_angle():
    """f_inf_rest = (1 + cos d_max)/2 with d_max the interpolated angle at the
    last table radius, here above pi/2 (so not the pi/2 floor)."""
    lp = affine_lp_table()
    sp, p = _lp_sys(lp, 0.5, 7.7, 0.0)
    D = lp_affine_value(LP_DEL, lp, 0.5, 7.7)
    d_max = D + LP_DEL_S * np.log(1000.0)
    assert d_max > np.pi / 2 + 0.05
    assert abs(sp["refl"][4] - 0.5 * (1 + np.cos(d_max))) < 1e-7
    assert abs(sp["refl"][4] - 0.5) > 1e-2


def test_lp_relline_flux_scales_with_the_interpolated_intensity():
    """relline_lp with the physical normalisation on the affine table is
    A(a, h) times its value on the unit table, bin by bin."""
    rel, _, _ = tables()
    e = log_grid(600, 0.1, 10.0)
    for a, h, _ in LP_POINTS[:3]:
        par = [6.4, h, a, 30.0, -1.0, 400.0, 0.0, 0.0, 2.0, 0.0]
        fA, st = oracle.eval_model(oracle.MOD_RELLINELP, oracle.Tables(rel, affine_lp_table()), par, e,
                                   phys_norm=True)
        f1, st1 = oracle.eval_model(oracle.MOD_RELLINELP, oracle.Tables(rel, affine_lp_table(unit=True)), par, e,
                                    phys_norm=True)
        assert st == 0 and st1 == 0
        sel = f1 > 1e-10 * f1.max()
        A = lp_affine_value(LP_INT, affine_lp_table(), a, h)
        assert np.max(np.abs(fA[sel] / f1[sel] / A - 1)) < 1e-6
