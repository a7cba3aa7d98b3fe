"""Pins the FP64 oracle: the reference's own known-answer tests (golden
fixture tests/golden/reference_kats.json), closed forms, brute force on small
inputs, and the invariants the reference's tests assert.  CPU only."""
import json
import os

import numpy as np
import pytest

import oracle
from tests.helpers import RELLINE_CFG0, log_grid, reflection_spectrum, tables

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def test_rebin_kat():
    k = GOLD["rebin_spectrum"]
    out = oracle.rebin(k["ener"], k["ener0"], k["val0"])
    assert np.max(np.abs(out - np.array(k["expected"]))) < k["tol"]


def test_rebin_mean_flux_kat():
    k = GOLD["rebin_mean_flux"]
    out = oracle.rebin_mean_flux(k["ener"], k["ener0"], k["val0"])
    assert np.max(np.abs(out - np.array(k["expected"]))) < k["tol"]


def test_interp_kats():
    k = GOLD["interp_lin_1d"]
    assert abs(oracle.lib.orc_interp_lin_1d(k["ifac"], k["rlo"], k["rhi"]) - k["expected"]) < k["tol"]
    k = GOLD["interp_lin_2d"]
    args = (k["ifac1"], k["ifac2"], k["r11"], k["r12"], k["r21"], k["r22"])
    assert abs(oracle.lib.orc_interp_lin_2d(*args) - k["expected"]) < k["tol"]
    assert abs(oracle.lib.orc_interp_lin_2d_float(*args) - k["expected"]) < k["tol"]


def test_kerr_closed_forms():
    k = GOLD["kerr_isco"]
    for a, r, tol in zip(k["a"], k["rms"], k["tol"]):
        assert abs(oracle.lib.orc_kerr_rms(a) - r) < tol, (a, oracle.lib.orc_kerr_rms(a))
    assert oracle.lib.orc_kerr_rplus(0.0) == 2.0
    assert abs(oracle.lib.orc_kerr_rplus(0.998) - (1 + np.sqrt(1 - 0.998 ** 2))) < 1e-15
    # ISCO is monotonically decreasing and concave in a: table chords stay below it
    a = np.linspace(-0.998, 0.998, 401)
    r = np.array([oracle.lib.orc_kerr_rms(x) for x in a])
    assert np.all(np.diff(r) < 0)
    assert np.all(np.diff(r, 2) < 1e-12)


def test_log_and_fine_grids_closed_form():
    e = oracle.log_grid(4097, 0.00035, 2000.0)
    assert abs(e[0] - 0.00035) < 1e-15 and abs(e[-1] - 2000.0) / 2000.0 < 1e-12
    assert np.allclose(np.diff(np.log(e)), np.log(2000.0 / 0.00035) / 4096, rtol=1e-9)
    re = oracle.fine_radial_grid(1.2369, 400.0, 1000)
    assert abs(re[0] - 400.0) / 400 < 1e-13 and abs(re[-1] - 1.2369) / 1.2369 < 1e-13
    assert np.allclose(np.diff(re ** -0.5), (1.2369 ** -0.5 - 400 ** -0.5) / 999, rtol=1e-9)


def test_energy_grid_shifts():
    """lineE and redshift act as E/lineE and E(1+z) on the grid
    (test/unit/test-cppspectrum.cpp:57-77)."""
    rel, lp, T = tables()
    e = log_grid(600, 0.1, 10.0)
    p = list(RELLINE_CFG0)
    f1, st = oracle.eval_model(oracle.MOD_RELLINE, T, p, e)
    assert st == 0
    p2 = list(p)
    p2[0] = 1.0
    f2, _ = oracle.eval_model(oracle.MOD_RELLINE, T, p2, e / 6.4)
    assert np.array_equal(f1, f2)
    p3 = list(p)
    p3[8] = 0.5
    f3, _ = oracle.eval_model(oracle.MOD_RELLINE, T, p3, e)
    f4, _ = oracle.eval_model(oracle.MOD_RELLINE, T, p, e * 1.5)
    assert np.array_equal(f3, f4)


def test_fft_matches_dft():
    rng = np.random.default_rng(0)
    x = rng.standard_normal(256) + 1j * rng.standard_normal(256)
    assert np.max(np.abs(oracle.fft(x) - np.fft.fft(x))) < 1e-11
    assert np.max(np.abs(oracle.fft(x, inverse=True) - np.fft.ifft(x) * 256)) < 1e-11


def test_fft_convolution_brute_force():
    """fftw_conv_spectrum (src/Relbase.cpp:119-188) equals the direct circular
    convolution of the E/dE-weighted spectra with the 1 keV rotation."""
    rng = np.random.default_rng(1)
    n = 256
    e = log_grid(n, 0.05, 20.0)
    fx = rng.uniform(0.1, 1.0, n)
    fr = rng.uniform(0.0, 1.0, n)
    c = 0.5 * (e[1:] + e[:-1]) / np.diff(e)
    i1 = oracle.lib.orc_binary_search(e.ctypes.data_as(oracle._D), n + 1, 1.0)
    x = fx * c
    y = np.zeros(n)
    for i in range(n):
        y[(i - i1 + n) % n] = fr[i] * c[i]
    direct = np.array([sum(x[m] * y[(k - m) % n] for m in range(n)) for k in range(n)]) * n
    out = oracle.fft_conv(e, fx, fr, normalized=False)
    assert np.max(np.abs(out * c - direct)) / np.max(np.abs(direct)) < 1e-12
    # delta input reproduces the rotated, re-weighted kernel
    fd = np.zeros(n)
    fd[i1] = 1.0
    out = oracle.fft_conv(e, fd, fr, normalized=False)
    expect = np.array([n * c[i1] * y[(k - i1) % n] / c[k] for k in range(n)])
    assert np.max(np.abs(out - expect)) / np.max(np.abs(expect)) < 1e-12


def test_conv_normalisation_invariants():
    """relline on the convolution grid sums to 1; the normalised convolution keeps
    the reflection flux in 0.01-1000 keV; a corrupted profile does not
    (test/unit/test-stdfunctions.cpp:253-298)."""
    k = GOLD["fft_normalisation"]
    rel, lp, T = tables()
    cg = oracle.log_grid(4097, 0.00035, 2000.0)
    p, st = oracle.rel_params(oracle.MOD_RELLINE, [1.0] + RELLINE_CFG0[1:])
    prof, _, st = oracle.relbase(T, p, cg)
    assert st == 0
    assert abs(prof.sum() - k["relline_sum"]) < k["relline_sum_tol"]
    e_ref, refl = reflection_spectrum(n_incl=1)
    xill = oracle.rebin(cg, e_ref, refl[0])
    band = (cg[:-1] >= k["band_emin"]) & (cg[1:] <= k["band_emax"])
    out = oracle.fft_conv(cg, xill, prof)
    assert abs(out[band].sum() - xill[band].sum()) < k["band_sum_tol"] * max(1.0, xill[band].sum())
    bad = prof.copy()
    bad[k["corrupt_bin"]] = k["corrupt_value"]
    out = oracle.fft_conv(cg, xill, bad)
    assert abs(out[band].sum() - xill[band].sum()) > k["band_sum_tol"]


def test_relconv_outside_band_is_zero():
    """tests-execmodel.cpp:211-235: relconv outside 0.01-1000 keV returns zeros,
    inside it keeps the normalisation."""
    k = GOLD["relconv_outside_band"]
    rel, lp, T = tables()
    p = [3.0, 3.0, 15.0, 0.998, 30.0, -1.0, 400.0, 0.0]
    lo, hi, n = k["outside_grid"]
    e = log_grid(n, lo, hi)
    f = np.zeros(n)
    f[n // 2] = k["test_value"]
    out, st = oracle.eval_model(oracle.MOD_RELCONV, T, p, e, flux_in=f)
    assert st == 0 and abs(out.sum()) < k["outside_sum_tol"]
    lo, hi, n = k["inside_grid"]
    e = log_grid(n, lo, hi)
    f = np.zeros(n)
    f[n // 2] = k["test_value"]
    out, st = oracle.eval_model(oracle.MOD_RELCONV, T, p, e, flux_in=f)
    assert st == 0 and abs(out.sum() - k["test_value"]) < k["inside_sum_tol"] * k["test_value"]


def test_relconv_of_narrow_line_is_relline():
    """relconv applied to a narrow Gaussian line at 6.4 keV reproduces relline
    (test/e2e/tests/test_convolution_models.sl:7-33,72): both on 600 log bins
    over 0.2-10 keV, relconv rescaled to relline's sum, goodness over
    2 < E_lo < 6 keV below 8e-4; the convolution keeps the input's photon
    number within 1e-3 (test_convolution_models.sl:77)."""
    from math import erf, sqrt
    k = GOLD["relconv_vs_relline"]
    rel, lp, T = tables()
    e = log_grid(600, 0.2, 10.0)
    sig = 0.01
    cdf = np.array([0.5 * (1.0 + erf((x - 6.4) / (sqrt(2.0) * sig))) for x in e])
    f = np.diff(cdf)
    pc = [3.0, 3.0, 15.0, 0.998, 30.0, -1.0, 400.0, 0.0]
    conv, st = oracle.eval_model(oracle.MOD_RELCONV, T, pc, e, flux_in=f)
    assert st == 0
    line, st = oracle.eval_model(oracle.MOD_RELLINE, T, [6.4, 3.0, 3.0, 15.0, 0.998, 30.0, -1.0, 400.0, 0, 0], e)
    assert st == 0
    assert abs(conv.sum() / f.sum() - 1.0) < k["norm_tol"]
    conv = conv * line.sum() / conv.sum()
    sel = (e[:-1] > 2.0) & (e[:-1] < 6.0) & (conv != 0)
    goodness = np.sqrt(np.sum((line[sel] / conv[sel] - 1) ** 2)) / sel.sum()
    print("relconv vs relline: norm", conv.sum() / line.sum(), "goodness", goodness)
    assert goodness < k["goodness_limit"]


def test_relline_sums_to_one():
    rel, lp, T = tables()
    e = log_grid(3000, 0.1, 10.0)
    f, st = oracle.eval_model(oracle.MOD_RELLINE, T, RELLINE_CFG0, e)
    assert st == 0 and abs(f.sum() - 1.0) < 1e-12
    f, st = oracle.eval_model(oracle.MOD_RELLINELP, T, [6.4, 6.0, 0.998, 30.0, -1.0, 400.0, 0.0, 0.0, 2.0, 0.0], e)
    assert st == 0 and abs(f.sum() - 1.0) < 1e-12


def test_flat_transfer_function_gives_pi():
    """Closed form of the r/g* integration: with trff_1 + trff_2 = 1, a window
    of width 2e-3 around E = 1 and a normalised emissivity, the photon flux is
    sum_r weight emis * int_0^1 dg*/sqrt(g*(1-g*)) = pi/4 (edges, Romberg and
    the g* bracket logic all take part)."""
    from tests.helpers import flat_table
    rel = flat_table(0.999, 1.001)
    T = oracle.Tables(rel)
    e = np.linspace(0.998, 1.002, 4001)
    p = [1.0, 3.0, 3.0, 15.0, 0.5, 40.0, -1.0, 400.0, 0.0, 0.0]
    f, st = oracle.eval_model(oracle.MOD_RELLINE, T, p, e, phys_norm=True)
    assert st == 0
    assert abs(f.sum() / (np.pi / 4) - 1.0) < 2e-3


def test_angle_weighting_linearity():
    """Sum_mu w_mu (X_mu * R) == (sum_mu w_mu X_mu) * R for the unnormalised
    FFT convolution (the identity that allows per-angle convolution summed in
    frequency space)."""
    rel, lp, T = tables()
    cg = oracle.log_grid(1025, 0.01, 1000.0)
    e_ref, refl = reflection_spectrum(n_incl=5)
    rb = np.array([oracle.rebin(cg, e_ref, r) for r in refl])
    p, _ = oracle.rel_params(oracle.MOD_RELLINE, [1.0] + RELLINE_CFG0[1:])
    prof, _, _ = oracle.relbase(T, p, cg)
    w = np.array([0.1, 0.3, 0.25, 0.2, 0.15])
    lhs = sum(w[k] * oracle.fft_conv(cg, rb[k], prof, normalized=False) for k in range(5))
    rhs = oracle.fft_conv(cg, (w[:, None] * rb).sum(0), prof, normalized=False)
    assert np.max(np.abs(lhs - rhs)) / np.max(np.abs(rhs)) < 1e-12


def test_rzone_grid():
    g = oracle.rzone_grid(2.0, 400.0, 1, 6.0)
    assert list(g) == [2.0, 400.0]
    g = oracle.rzone_grid(2.0, 400.0, 10, 6.0)
    assert abs(g[0] - 2.0) < 1e-12 and abs(g[-1] - 400.0) < 1e-9 and np.all(np.diff(g) > 0)


def test_rebin_conserves_total():
    rng = np.random.default_rng(3)
    e0 = np.sort(rng.uniform(0.5, 50.0, 300))
    f0 = rng.uniform(0, 1, 299)
    e = log_grid(777, 0.1, 100.0)
    out = oracle.rebin(e, e0, f0)
    assert abs(out.sum() - f0.sum()) < 1e-10 * f0.sum()


def test_rebin_shared_top_edge_is_defined():
    """An output edge equal to the last input edge: the reference's walk reads
    past the input grid there (src/relutility.c:564); the overlap has zero
    width and the oracle defines the value as 0."""
    e0 = np.array([1.0, 2.0, 3.0])
    f0 = np.array([1.0, 1.0])
    out = oracle.rebin(np.array([3.0, 4.0, 5.0]), e0, f0)
    assert np.array_equal(out, [0.0, 0.0])
    out = oracle.rebin(np.array([2.5, 3.0]), e0, f0)
    assert abs(out[0] - 0.5) < 1e-15


@pytest.mark.parametrize("bad, status", [
    ([6.4, 3, 3, 15, 0.999, 30, -1, 400, 0, 0], 1),    # a > 0.9982
    ([6.4, 3, 3, 15, 0.5, 2.0, -1, 400, 0, 0], 1),      # incl < 3 deg
    ([6.4, 3, 3, 15, 0.5, 88.0, -1, 400, 0, 0], 1),     # incl > 87 deg
    ([6.4, 3, 3, 15, 0.5, 30, 50, 40, 0, 0], 1),        # Rin >= Rout
])
def test_parameter_rejections(bad, status):
    p, st = oracle.rel_params(oracle.MOD_RELLINE, bad)
    assert st == status


def test_parameter_clamps():
    p, st = oracle.rel_params(oracle.MOD_RELLINE, [6.4, 3, 3, 1000, 0.5, 30, -2, 2000, 0, 0])
    assert st == 0
    isco = oracle.lib.orc_kerr_rms(0.5)
    assert abs(p.rin - 2 * isco) < 1e-12 and p.rout == 1000.0 and p.rbr == 1000.0
    p, st = oracle.rel_params(oracle.MOD_RELLINE, [6.4, 3, 3, 1.0, 0.5, 30, 0.5, 400, 0, 0])
    assert st == 0 and p.rin == isco and p.rbr == isco
    p, st = oracle.rel_params(oracle.MOD_RELLINELP, [6.4, 1.0, 0.9, 30, -1, 400, 0, 0, 2, 0])
    assert st == 0 and abs(p.height - 1.1 * oracle.lib.orc_kerr_rplus(0.9)) < 1e-12
    p, st = oracle.rel_params(oracle.MOD_RELCONVLP, [-3.0, 1.5, 0.9, 30, -1, 400, 0, 2, 0])
    assert st == 0 and p.beta == 0.99 and abs(p.height - 3 * oracle.lib.orc_kerr_rplus(0.9)) < 1e-12
    p, st = oracle.rel_params(oracle.MOD_RELLINELP, [6.4, 6.0, 0.9, 30, -1, 400, 0, 0, 2, 1])
    assert st == 0 and p.return_rad == 1


def test_lp_closed_forms():
    for r, a, h in [(10.0, 0.5, 6.0), (3.0, 0.998, 3.0), (50.0, -0.5, 20.0)]:
        ut_d = (r * np.sqrt(r) + a) / (np.sqrt(r) * np.sqrt(r * r - 3 * r + 2 * a * np.sqrt(r)))
        ut_h = np.sqrt((h * h + a * a) / (h * h - 2 * h + a * a))
        assert abs(oracle.lib.orc_gi_potential_lp(r, a, h, 0.0, 1.0) - ut_d / ut_h) < 1e-13
    # Newtonian lamp post: (1/(2 pi h^2)) (1 + (r/h)^2)^(-3/2)
    h, r = 10.0, 30.0
    assert abs(oracle.lib.orc_lp_emissivity_newton(h, r) - (1 + (r / h) ** 2) ** -1.5 / (2 * np.pi * h * h)) < 1e-16


def test_work_counters_cfg0():
    """Work per cfg0 spectrum (used for the roofline's flop count)."""
    rel, lp, T = tables()
    e = log_grid(3000, 0.1, 10.0)
    f, st, s = oracle.eval_model(oracle.MOD_RELLINE, T, RELLINE_CFG0, e, stats=True)
    assert st == 0
    assert 1e5 < s["relb_func_calls"] < 1e7
    # every bin integral is a midpoint, a Romberg pair (two branches) or edge-only
    assert s["bin_integrals"] >= s["midpoint_calls"] + s["romberg_calls"] // 2


def test_sawtooth_transfer_function_per_bin_closed_form():
    """relb_func's swapped lerp weights (src/Relprofile.cpp:502-504), the g*
    grid (:104-116), the bracket search, E^3/sqrt(g*(1-g*)) and Romberg, bin by
    bin against quadrature of the closed form; the usual (unswapped) lerp
    would miss by percents."""
    from tests.helpers import sawtooth_bins, sawtooth_table
    rel = sawtooth_table()
    T = oracle.Tables(rel)
    e = np.linspace(0.998, 1.002, 4001)
    f, st = oracle.eval_model(oracle.MOD_RELLINE, T, [1.0, 3.0, 3.0, 15.0, 0.5, 40.0, -1.0, 400.0, 0.0, 0.0],
                              e, phys_norm=True)
    assert st == 0
    idx, ref = sawtooth_bins(e, rel)
    assert idx.size > 1500
    assert np.max(np.abs(f[idx] / ref - 1)) < 1e-6
    _, ref_u = sawtooth_bins(e, rel, swapped=False)
    assert np.max(np.abs(f[idx] / ref_u - 1)) > 1e-2  # the check can tell the two lerps apart


@pytest.mark.parametrize("a,incl", [(0.5, 40.0), (-0.63, 71.0), (0.93, 12.0)])
def test_affine_table_bilinear_closed_form(a, incl):
    """Bilinear (a, mu0) node interpolation with float factors
    (src/Relprofile.cpp:39-80, :171-177) on a field affine in a and mu0:
    per-bin flux = closed form of the flat window times the affine value."""
    from tests.helpers import affine_expected, affine_table, sawtooth_bins
    rel = affine_table()
    T = oracle.Tables(rel)
    e = np.linspace(0.998, 1.002, 4001)
    f, st = oracle.eval_model(oracle.MOD_RELLINE, T, [1.0, 3.0, 3.0, 15.0, a, incl, -1.0, 400.0, 0.0, 0.0],
                              e, phys_norm=True)
    assert st == 0
    k, (fa, fm) = affine_expected(rel, a, incl)
    idx, unit = sawtooth_bins(e, rel, c_rel=0.0)
    assert np.max(np.abs(f[idx] / (k * unit) - 1)) < 1e-6
    # exchanged a weights would be visible at this tolerance
    k_swap, _ = affine_expected(rel, a, incl, swap_a=True)
    assert abs(k_swap / k - 1) > 1e-4


@pytest.mark.parametrize("a,incl", [(0.5, 40.0), (-0.3, 60.0), (0.95, 20.0)])
def test_radial_mapping_closed_form(a, incl):
    """Node radii lerped in spin and fine-grid values lerped between node
    brackets (src/Relprofile.cpp:190-193, :249-300): a transfer function
    affine in the table radius gives, relative to the flat table, the weighted
    mean of (u + v re/1000) over the fine grid; a reversed lerp would not."""
    from tests.helpers import flat_table, radial_expected, radial_table
    e = np.linspace(0.998, 1.002, 4001)
    p = [1.0, 3.0, 3.0, 15.0, a, incl, -2.0, 400.0, 0.0, 0.0]  # Rin = 2 ISCO: clear of the ISCO snap
    fr, st = oracle.eval_model(oracle.MOD_RELLINE, oracle.Tables(radial_table()), p, e, phys_norm=True)
    ff, st2 = oracle.eval_model(oracle.MOD_RELLINE, oracle.Tables(flat_table(0.999, 1.001)), p, e,
                                phys_norm=True)
    assert st == 0 and st2 == 0
    rin = 2 * oracle.lib.orc_kerr_rms(a)
    got = fr.sum() / ff.sum()
    assert abs(got / radial_expected(rin, 400.0) - 1) < 1e-7
    rel = radial_table()
    r = rel["r"].reshape(rel["n_a"], rel["n_mu0"], rel["n_r"])[:, 0, :].astype(np.float64)
    ia = np.searchsorted(rel["a"], np.float32(a), side="right") - 1
    fa = float(np.float32((np.float32(a) - rel["a"][ia]) / (rel["a"][ia + 1] - rel["a"][ia])))
    node_r = fa * r[ia + 1] + (1 - fa) * r[ia]
    assert abs(got / radial_expected(rin, 400.0, node_r=node_r) - 1) > 1e-5


def test_outer_radius_snap_closed_form():
    """The 1000 rg snap (src/Relprofile.cpp:235-238): the table's outermost node
    radius 999 is moved to 1000, so on a radius-affine transfer function with a
    flat emissivity (q = 0, outer radii dominate) the fine values in the outer
    node interval follow the ramp from (r_1, u + v r_1/1000) to (1000,
    u + v 999/1000) instead of u + v re/1000."""
    from tests.helpers import flat_table, radial_table
    e = np.linspace(0.998, 1.002, 4001)
    a, incl = 0.5, 40.0
    p = [1.0, 0.0, 0.0, 15.0, a, incl, -2.0, 1000.0, 0.0, 0.0]
    rel = radial_table()
    fr, st = oracle.eval_model(oracle.MOD_RELLINE, oracle.Tables(rel), p, e, phys_norm=True)
    ff, st2 = oracle.eval_model(oracle.MOD_RELLINE, oracle.Tables(flat_table(0.999, 1.001)), p, e, phys_norm=True)
    assert st == 0 and st2 == 0
    got = fr.sum() / ff.sum()
    rin = 2 * oracle.lib.orc_kerr_rms(a)
    re = oracle.fine_radial_grid(rin, 1000.0, 1000)
    dr = np.empty_like(re)
    dr[0] = 0.5 * (re[0] - re[1])
    dr[-1] = 0.5 * (re[-2] - re[-1])
    dr[1:-1] = 0.5 * (re[:-2] - re[2:])
    w = re * dr  # q = 0
    r = rel["r"].reshape(rel["n_a"], rel["n_mu0"], rel["n_r"])[:, 0, :].astype(np.float64)
    ia = np.searchsorted(rel["a"], np.float32(a), side="right") - 1
    fa = float(np.float32((np.float32(a) - rel["a"][ia]) / (rel["a"][ia + 1] - rel["a"][ia])))
    r0, r1 = fa * r[ia + 1, 0] + (1 - fa) * r[ia, 0], fa * r[ia + 1, 1] + (1 - fa) * r[ia, 1]
    assert r0 < 1000.0 < r0 * 1.01  # the snap applies
    u, v = 0.01, 1.0
    x = u + v * re / 1000.0
    outer = re > r1
    ramp = (re[outer] - r1) / (1000.0 - r1)
    x_snap = x.copy()
    x_snap[outer] = (1 - ramp) * (u + v * r1 / 1000.0) + ramp * (u + v * r0 / 1000.0)
    expect_snap = np.sum(w * x_snap) / np.sum(w)
    expect_plain = np.sum(w * x) / np.sum(w)
    assert abs(got / expect_snap - 1) < 1e-7
    assert abs(got / expect_plain - 1) > 1e-5


def test_lp_radius_factor_switch_closed_form():
    """rebin_emisprofile_on_radial_grid (src/Rellp.cpp:120-175) interpolates in
    log r where the lower node's emission angle exceeds 75 deg and linearly in
    r elsewhere (get_ipol_factor_radius, src/relutility.c:407-417), with intensity log-
    interpolated.  On _switch_lp_table each region's fields are reproduced
    exactly by its own factor and missed by the other one (by 1e-5 .. 1e-3);
    the node interval that straddles r_s follows the lower node's (log) factor:
    a power law through its two intensity nodes."""
    r_s = 10.0
    from tests.helpers import switch_fields, switch_lp_table
    rel, lp = switch_lp_table(r_s)
    T = oracle.Tables(rel, lp)
    a = float(lp["a"][11])  # a table spin: no spin interpolation
    gamma, h = 2.0, 6.0
    rp, st = oracle.rel_params(oracle.MOD_RELLINELP, [6.4, h, a, 30.0, -1.0, 100.0, 0.0, 0.0, gamma, 0.0])
    assert st == 0
    sp, st = oracle.system_parameters(T, rp)
    assert st == 0
    re = sp["re"]
    ut_d = (re * np.sqrt(re) + a) / (np.sqrt(re) * np.sqrt(re * re - 3 * re + 2 * a * np.sqrt(re)))
    ut_h = np.sqrt((h * h + a * a) / (h * h - 2 * h + a * a))
    emis = sp["emis"] / (ut_d / ut_h) ** gamma  # beta = 0: the boost is g^gamma
    rad = lp["rad"].reshape(lp["n_a"], lp["n_r"])[11].astype(np.float64)
    k = np.searchsorted(rad, re, side="right") - 1
    lo_node = rad[k] < r_s
    straddle = lo_node & (rad[k + 1] >= r_s)
    inside = ~straddle
    i_x, d_x, di_x = switch_fields(re, re < r_s)
    assert inside.sum() > 500 and straddle.sum() > 5 and (re[inside] < r_s).sum() > 100
    assert np.max(np.abs(emis[inside] / i_x[inside] - 1)) < 2e-6
    assert np.max(np.abs(sp["del_emit"][inside] - d_x[inside])) < 1e-6
    assert np.max(np.abs(sp["del_inc"][inside] - di_x[inside])) < 1e-6
    # the other factor in each region misses
    f_lin = (re - rad[k]) / (rad[k + 1] - rad[k])
    i_lo, d_lo, _ = switch_fields(rad[k], rad[k] < r_s)
    i_hi, d_hi, _ = switch_fields(rad[k + 1], rad[k + 1] < r_s)
    f_log = np.log(re / rad[k]) / np.log(rad[k + 1] / rad[k])
    f_wrong = np.where(lo_node, f_lin, f_log)
    d_wrong = (1 - f_wrong) * d_lo + f_wrong * d_hi
    i_wrong = i_lo * (i_hi / i_lo) ** f_wrong
    miss_d = np.abs(d_wrong - d_x)[inside]
    miss_i = np.abs(i_wrong / i_x - 1)[inside]
    mid = inside & (f_lin > 0.3) & (f_lin < 0.7)
    assert np.min(np.abs(d_wrong - d_x)[mid]) > 3e-6 and np.max(miss_d) > 2e-5
    assert np.min(np.abs(i_wrong / i_x - 1)[mid]) > 5e-6 and np.max(miss_i) > 1e-4
    # straddling interval: lower node > 75 deg -> power law through the nodes
    pl = i_lo * (re / rad[k]) ** (np.log(i_hi / i_lo) / np.log(rad[k + 1] / rad[k]))
    assert np.max(np.abs(emis[straddle] / pl[straddle] - 1)) < 2e-6
    lin = i_lo * (i_hi / i_lo) ** f_lin
    assert np.max(np.abs(lin[straddle] / pl[straddle] - 1)) > 5e-3


def test_isco_snap_closed_form():
    """The ISCO snap (src/Relprofile.cpp:195-200): an innermost node radius
    within 1e-3 above the ISCO is moved onto it.  Table: innermost radius
    ISCO (1 + 5e-4) at every spin node, transfer function equal to the table
    radius (trff_1 = trff_2 = r/2).  At a table spin the fine values equal re
    except in the innermost node interval, which ramps from (r_1, r_1) to
    (ISCO, r_0 (1 + 5e-4)); without the snap the mean would be that of re."""
    from tests.helpers import flat_table, isco_snap_table
    rel = isco_snap_table()
    ia = int(np.searchsorted(rel["a"], 0.5))
    a = float(rel["a"][ia])
    rms = oracle.lib.orc_kerr_rms(a)
    nodes = rel["r"].reshape(rel["n_a"], rel["n_mu0"], rel["n_r"])[ia, 0].astype(np.float64)
    assert nodes[-1] > rms and (nodes[-1] - rms) / nodes[-1] < 1e-3
    e = np.linspace(0.998, 1.002, 4001)
    p = [1.0, 3.0, 3.0, 15.0, a, 40.0, -1.0, 20.0, 0.0, 0.0]
    fr, st = oracle.eval_model(oracle.MOD_RELLINE, oracle.Tables(rel), p, e, phys_norm=True)
    ff, st2 = oracle.eval_model(oracle.MOD_RELLINE, oracle.Tables(flat_table(0.999, 1.001)), p, e, phys_norm=True)
    assert st == 0 and st2 == 0
    got = fr.sum() / ff.sum()
    re = oracle.fine_radial_grid(rms, 20.0, 1000)
    dr = np.empty_like(re)
    dr[0] = 0.5 * (re[0] - re[1])
    dr[-1] = 0.5 * (re[-2] - re[-1])
    dr[1:-1] = 0.5 * (re[:-2] - re[2:])
    w = re * dr * re ** -3.0
    x = re.copy()
    last = re < nodes[-2]
    ramp = (nodes[-2] - re[last]) / (nodes[-2] - rms)
    x[last] = (1 - ramp) * nodes[-2] + ramp * nodes[-1]
    assert last.sum() > 10
    assert abs(got / (np.sum(w * x) / np.sum(w)) - 1) < 1e-7
    assert abs(got / (np.sum(w * re) / np.sum(w)) - 1) > 1e-6
