"""The xillver table and the relxill model on the CPU: the oracle's table
interpolation (src/xilltable.c:297-324, :478-566, :812-876, :1054-1181), the
primary continuum and its normalisation (src/Xillspec.cpp:179-241), the relxill
assembly (src/Relxill.cpp:289-482, src/PrimarySource.cpp:66-125) pinned by
closed forms, and the library's xillver table loader and synthetic table.
No GPU needed."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import oracle
import relxill_amd as rx
from tests.helpers import (affine_xill_expected, affine_xilltable, log_grid, rel_err, tables, write_fits_xilltable,
                           xill_axes)

CASES = [(2.1, 1.3, 3.1, 250.0), (1.73, 6.2, 0.4, 17.0), (3.05, 0.71, 4.45, 612.0),
         # below and above every axis: the values are clamped into the table
         (0.5, 0.1, -1.0, 2.0), (3.9, 20.0, 6.0, 2000.0),
         # exactly on nodes, including the last one
         (1.8, 2.236068, 2.35, 170.99759), (3.4, 10.0, 4.7, 1000.0)]


@pytest.fixture(scope="module")
def affine():
    x, g = affine_xilltable()
    return x, g, oracle.Xill(x)


@pytest.mark.parametrize("gam,afe,lxi,ect", CASES)
def test_interpolation_reproduces_an_affine_table(affine, gam, afe, lxi, ect):
    """Stored spectra 10^logXi g(E) (c0 + c1 Gamma + c2 A_Fe + c3 logXi + c4
    Ecut + c5 k): the division by 10^logXi and the 16-corner interpolation
    give back the affine value at the (clamped) parameters for every
    inclination k.  A dropped renormalisation, swapped axes or weights fail."""
    x, g, xt = affine
    e, flu = oracle.xill_spectra(xt, oracle.xill_param(gam, afe, lxi, ect))
    want = affine_xill_expected(x, g, gam, afe, lxi, ect)
    assert np.abs(flu / want - 1).max() < 2e-6
    # the spectrum's grid: ENERG_LO and the last ENERG_HI (src/xilltable.c:1104-1108)
    assert np.array_equal(e[:-1], x["elo"].astype(np.float64)) and e[-1] == np.float64(x["ehi"][-1])


def test_corner_indices_and_factors(affine):
    """binary_search_float with the index kept in [0, n-2]; the factor is
    formed in float after clamping the value into the table; a value on a
    node gives factor 0 (1 on the last node); Ecut at or beyond the ends gives
    0 / 1 (ensure_ecut_within_boundarys)."""
    x, _, xt = affine
    v = x["val"]
    ind, fac = oracle.xill_corner(xt, oracle.xill_param(float(v[0][1]), float(v[1][0]), float(v[2][-1]),
                                                        float(v[3][2])))
    assert list(ind[:4]) == [1, 0, len(v[2]) - 2, 2]
    assert list(fac[:4]) == [0.0, 0.0, 1.0, 0.0]
    mid = 0.5 * (float(v[0][1]) + float(v[0][2]))
    ind, fac = oracle.xill_corner(xt, oracle.xill_param(mid, 0.01, 99.0, 1e5))
    assert ind[0] == 1 and fac[0] == pytest.approx(0.5, abs=1e-6)
    assert (ind[1], fac[1]) == (0, 0.0) and (ind[2], fac[2]) == (len(v[2]) - 2, 1.0) and fac[3] == 1.0
    # the factor is a float quotient: it carries float rounding, not double
    gam = 2.1
    ind, fac = oracle.xill_corner(xt, oracle.xill_param(gam, 1.0, 1.0, 100.0))
    lo, hi = v[0][ind[0]], v[0][ind[0] + 1]
    assert fac[0] == float(np.float32((np.float32(gam) - lo) / (hi - lo)))


def test_density_axis_renormalisation():
    """A table with a density axis instead of Ecut (as xillverD): rows are
    divided by 10^logXi and 10^(logN-15) of their own indices
    (normalizeXillverSpecLogxiDensity, src/xilltable.c:488-511); the model's
    default density 15 is then interpolated like any other parameter."""
    x, g = affine_xilltable()
    ax = xill_axes()
    dens = np.array([15.0, 16.0, 17.5, 19.0], dtype=np.float32)
    G, A, L, D, K = np.meshgrid(ax[0], ax[1], ax[2], dens, np.arange(len(ax[4])), indexing="ij")
    lin = 0.7 + 0.21 * G + 0.037 * A + 0.083 * L + 0.05 * D + 0.059 * K
    x = dict(x, pid=[0, 1, 2, 4, 7], val=[x["val"][0], x["val"][1], x["val"][2], dens, x["val"][4]],
             spec=((10.0 ** L * 10.0 ** (D - 15) * lin)[..., None] * g).reshape(-1, g.size).astype(np.float32))
    xt = oracle.Xill(x)
    for d in (15.0, 16.4, 30.0):
        e, flu = oracle.xill_spectra(xt, oracle.xill_param(2.1, 1.3, 3.1, 300.0, dens=d))
        dd = min(d, 19.0)
        want = (0.7 + 0.21 * 2.1 + 0.037 * 1.3 + 0.083 * 3.1 + 0.05 * dd + 0.059 * np.arange(6))[:, None] * g
        assert np.abs(flu / want - 1).max() < 2e-6, d


def test_angle_weighting_is_the_weighted_sum():
    """calc_xillver_angdep (src/Xillspec.cpp:557-573) on a table whose angle
    rows are unit vectors: the result is the distribution itself."""
    n_incl, n = 6, 6
    flu = np.eye(n_incl, n)
    dist = np.array([0.1, 0.25, 0.0, 0.4, 0.2, 0.05])
    out = np.zeros(n)
    oracle.lib.orc_xill_angdep(oracle._dp(out), oracle._dp(oracle._arr(flu.reshape(-1))), oracle._dp(dist),
                               n_incl, n)
    assert np.array_equal(out, dist)


@pytest.mark.parametrize("gam,ect", [(2.0, 300.0), (1.4, 50.0), (3.0, 20.0), (1.0, 5.0)])
def test_primary_normalisation_quadrature(gam, ect):
    """1 / calcNormWrtXillverTableSpec of exp(1/Ecut) E^-Gamma exp(-E/Ecut) on
    500 log bins over 0.1-1000 keV: the midpoint sum of E times the photon flux
    is the integral of the energy flux to 2e-4; a dropped exp(1/Ecut), keV->erg
    factor, 4 pi or end bin misses by more."""
    g = log_grid(500, 0.1, 1000.0)
    assert g[0] >= 0.1  # every bin is inside the band on this grid
    integral = quad(lambda e: math.exp(1 / ect) * e ** (1 - gam) * math.exp(-e / ect), g[0], g[-1],
                    limit=400, epsabs=0, epsrel=1e-12)[0]
    want = 1.0 / (integral * 1e20 * 1.602177e-09 / (1e15 / 4 / math.pi))
    assert oracle.primary_norm_factor(gam, ect) / want - 1 == pytest.approx(0.0, abs=2e-4)


@pytest.fixture(scope="module")
def relx():
    _, _, T = tables()
    rx.xilltable_synthetic(5, 5, 4, 6, 5, 10, 600)
    x = rx.xilltable_view()
    return T, oracle.Xill(x)


RELXILL_P = [3.0, 3.0, 15.0, 0.998, 30.0, -1.0, 400.0, 0.0, 2.0, 3.1, 1.0, 300.0, 3.0]


@pytest.mark.parametrize("gam,ect,z", [(2.0, 300.0, 0.0), (1.6, 40.0, 0.3), (2.8, 15.0, 1.5)])
def test_relxill_primary_only_closed_form(relx, gam, ect, z):
    """refl_frac = 0: the reflection is multiplied by 0 and the output is the
    normalised cut-off power law on the redshifted caller grid, i.e. the
    integral of nf exp(1/Ecut) E^-Gamma exp(-E/Ecut) over [E_j, E_j+1](1+z)
    (the 4096-bin midpoint sum and the linear split of partial bins agree to
    1e-3)."""
    T, xt = relx
    p = list(RELXILL_P)
    p[7], p[8], p[11], p[12] = z, gam, ect, 0.0
    e = log_grid(200, 0.1, 100.0)
    y, _, st = oracle.relxill(T, xt, p, e)
    assert st == 0
    nf = oracle.primary_norm_factor(gam, ect)
    es = e * (1 + z)
    want = np.array([nf * quad(lambda x: math.exp(1 / ect) * x ** -gam * math.exp(-x / ect), a, b,
                                epsrel=1e-10)[0] for a, b in zip(es[:-1], es[1:])])
    assert rel_err(y, want) < 1e-3


def test_relxill_linear_in_refl_frac(relx):
    """add_primary_spectrum for the broken power law: reflection times
    |refl_frac|, the primary added only for refl_frac >= 0."""
    T, xt = relx
    e = log_grid(300, 0.1, 100.0)
    run = lambda rf: oracle.relxill(T, xt, RELXILL_P[:12] + [rf], e)[0]
    r1, r2, p0, p3 = run(-1.0), run(-2.0), run(0.0), run(3.0)
    assert r1.max() > 0 and p0.max() > 0
    assert rel_err(r2, 2 * r1) < 1e-14
    assert rel_err(p3, 3 * r1 + p0) < 1e-13


def test_relxill_reflection_is_the_angle_resolved_convolution(relx):
    """relxill with refl_frac = -1 is the relxill-style convolution of the
    interpolated per-angle xillver spectra (the same profile, angle weights,
    rebin and FFT normalisation), with the table's energy grid."""
    T, xt = relx
    e = log_grid(300, 0.1, 100.0)
    for p in (RELXILL_P[:12] + [-1.0], [2.0, 4.0, 8.0, -0.5, 70.0, 3.0, 300.0, 0.0, 1.7, 1.2, 3.3, 90.0, -1.0]):
        y, dist, st = oracle.relxill(T, xt, p, e)
        assert st == 0
        xp = oracle.xill_param(p[8], p[10], p[9], p[11])
        ex, flu = oracle.xill_spectra(xt, xp)
        conv_p = p[:7] + [0.0]
        cg = log_grid(4096, 0.00035, 2000.0)
        ref, dist_ref, st = oracle.relxill_conv(T, conv_p, cg, ex, flu, e)
        assert st == 0
        assert np.array_equal(dist, dist_ref)
        assert rel_err(y, ref) < 1e-13


def test_library_xilltable_loader_round_trip(tmp_path):
    """rx_xilltable_load_fits reads what the OGIP writer wrote: axes, ids,
    energy bins and every spectrum, unchanged."""
    x, _ = affine_xilltable(n_ener=120)
    write_fits_xilltable(tmp_path / "xillver-a-Ec5.fits", x)
    try:
        assert rx.lib.rx_xilltable_load_fits(str(tmp_path / "xillver-a-Ec5.fits").encode()) == 0, \
            rx.lib.rx_last_error()
        v = rx.xilltable_view()
        assert v["pid"] == [0, 1, 2, 3, 7] and v["n_val"] == x["n_val"] and v["n_ener"] == 120
        for k in range(5):
            assert np.array_equal(v["val"][k], x["val"][k])
        assert np.array_equal(v["elo"], x["elo"]) and np.array_equal(v["ehi"], x["ehi"])
        assert np.array_equal(v["spec"], x["spec"])
        # malformed: an unknown parameter name, and a table missing spectra rows
        bad = dict(x, pid=[0, 1, 2, 3, 5])
        from tests import helpers
        helpers.XILL_NAMES[5] = "Foo"
        write_fits_xilltable(tmp_path / "bad.fits", bad)
        assert rx.lib.rx_xilltable_load_fits(str(tmp_path / "bad.fits").encode()) == rx.RX_ERR_TABLE
        assert "unknown parameter Foo" in rx.lib.rx_last_error().decode()
        short = dict(x, spec=x["spec"][:-7])
        write_fits_xilltable(tmp_path / "short.fits", short)
        assert rx.lib.rx_xilltable_load_fits(str(tmp_path / "short.fits").encode()) == rx.RX_ERR_TABLE
        assert "fewer rows" in rx.lib.rx_last_error().decode()
        # failed loads leave the resident table untouched
        assert rx.xilltable_view()["n_ener"] == 120
    finally:
        rx.xilltable_synthetic()


def test_synthetic_xilltable_shape():
    """The default synthetic table: Gamma, A_Fe, logXi, Ecut, Incl on the
    xillver-a-Ec5 ranges, cos-uniform inclination bins, positive spectra
    stored times 10^logXi."""
    rx.xilltable_synthetic(5, 5, 3, 4, 3, 10, 400)
    try:
        v = rx.xilltable_view()
        assert v["pid"] == [0, 1, 2, 3, 7]
        assert v["val"][0][0] == 1.0 and v["val"][0][-1] == np.float32(3.4)
        assert v["val"][2][0] == 0.0 and v["val"][2][-1] == np.float32(4.7)
        mu = np.cos(np.radians(v["val"][4].astype(np.float64)))
        assert np.allclose(mu, 1 - (np.arange(10) + 0.5) / 10, atol=1e-6)
        assert v["elo"][0] == np.float32(0.07) and v["ehi"][-1] == np.float32(1000.0)
        s = v["spec"].reshape(5, 3, 4, 3, 10, 400)
        assert (s >= 0).all() and s[..., :200].min() > 0
        # stored times 10^logXi: the renormalised rows of two logXi agree in scale
        r0 = s[2, 1, 0, 1, 5].astype(np.float64)
        r2 = s[2, 1, 2, 1, 5].astype(np.float64) / 10 ** float(v["val"][2][2])
        assert 0.2 < np.median(r2[300:] / r0[300:]) < 5
    finally:
        rx.xilltable_synthetic()


# ---- relxilllp: lamp post, 10 radial zones ----

RELXILLLP_P = [6.0, 0.0, 0.9, 30.0, -1.0, 400.0, 0.0, 2.0, 3.1, 1.0, 300.0, 1.0, 0.0, 0.0]


@pytest.fixture(scope="module")
def relxlp():
    from tests.helpers import tables as _tables
    _, _, T = _tables()
    rx.xilltable_synthetic(5, 5, 4, 6, 5, 10, 600)
    return T, oracle.Xill(rx.xilltable_view())


def _ut_disk(r, a):
    return (r * math.sqrt(r) + a) / (math.sqrt(r) * math.sqrt(r * r - 3 * r + 2 * a * math.sqrt(r)))


def _norm_integral(gam, ect):
    return quad(lambda e: math.exp(1 / ect) * e ** (1 - gam) * math.exp(-e / ect), 0.1, 1000.0,
                limit=400, epsabs=0, epsrel=1e-12)[0]


@pytest.mark.parametrize("h,a,beta,incl", [(6.0, 0.9, 0.0, 30.0), (3.0, 0.998, 0.0, 60.0), (20.0, -0.5, 0.3, 45.0)])
def test_relxilllp_zones_closed_forms(relxlp, h, a, beta, incl):
    """The zones (RadialGrid, IonGradient): edges from Rin to Rout, log-spaced
    below the source height and uniform in 1/r above; the source -> disk energy
    shift of a static source is u^t(disk, r) / u^t(axis, h) at the zone centre;
    the observer shift sqrt(1 - 2h/(h^2+a^2)) times the Doppler factor
    sqrt(1-beta^2)/(1 - beta cos i); the zone's Ecut is the observed Ecut
    over the observer shift times the disk shift; the normalisation change is
    the ratio of the incident energy-flux integrals (source over zone), 2e-4."""
    T, _ = relxlp
    p = list(RELXILLLP_P)
    p[0], p[1], p[2], p[3] = h, beta, a, incl
    z, st = oracle.relxill_zones(oracle.MOD_RELXILLLP, T, p)
    assert st == 0
    rin = float(oracle.lib.orc_kerr_rms(a))
    edges = np.append(z[:, 0], z[-1, 1])
    assert edges[0] == rin and edges[-1] == pytest.approx(400.0, rel=1e-14) and np.all(np.diff(edges) > 0)
    # log grid up to its last edge at or below h, then uniform in 1/r
    logg = rin * (400.0 / rin) ** (np.arange(11) / 10)
    indr = int(np.searchsorted(logg, h, side="right") - 1) if h > rin else 0
    assert np.allclose(edges[:indr + 1], logg[:indr + 1], rtol=1e-12)
    inv = 1.0 / edges[indr:]
    assert np.allclose(np.diff(inv), np.diff(inv)[0], rtol=1e-9)
    g_so = math.sqrt(1 - 2 * h / (h * h + a * a))
    if beta >= 1e-4:
        g_so *= math.sqrt(1 - beta * beta) / (1 - beta * math.cos(math.radians(incl)))
    assert oracle.energy_shift_source_obs(oracle.MOD_RELXILLLP, p)[0] == pytest.approx(g_so, rel=1e-13)
    ut_h = math.sqrt((h * h + a * a) / (h * h - 2 * h + a * a))
    ect_src = p[10] / g_so
    for iz in range(10):
        rm = 0.5 * (z[iz, 0] + z[iz, 1])
        if beta == 0.0:
            assert z[iz, 3] == pytest.approx(_ut_disk(rm, a) / ut_h, rel=1e-12)
        assert z[iz, 4] == pytest.approx(ect_src * z[iz, 3], rel=1e-13)
        want = _norm_integral(2.0, ect_src) / _norm_integral(2.0, z[iz, 4])
        assert z[iz, 5] / want - 1 == pytest.approx(0.0, abs=2e-4)


def test_relxilllp_zone_profiles_partition_the_profile(relxlp):
    """The zones' profiles add up to the one-zone profile (the lamp post is not
    renormalised), and each zone's profile lies inside the energy windows of
    the fine radii between its edges."""
    T, _ = relxlp
    p, st = oracle.rel_params(oracle.MOD_RELXILLLP, RELXILLLP_P)
    assert st == 0
    cg = log_grid(4096, 0.00035, 2000.0)
    f10, d10, st = oracle.relbase(T, p, cg, n_incl=10)
    assert st == 0
    p.num_zones = 1
    f1, _, st = oracle.relbase(T, p, cg, n_incl=10)
    f10 = f10.reshape(10, -1)
    assert rel_err(f10.sum(axis=0), f1) < 1e-13
    p.num_zones = 10
    sp, st = oracle.system_parameters(T, p)
    z, _ = oracle.relxill_zones(oracle.MOD_RELXILLLP, T, RELXILLLP_P)
    emid = 0.5 * (cg[1:] + cg[:-1])
    for iz in range(10):
        sel = (sp["re"] >= z[iz, 0]) & (sp["re"] < z[iz, 1]) if iz < 9 else (sp["re"] >= z[iz, 0])
        lo, hi = sp["gmin"][sel].min(), sp["gmax"][sel].max()
        outside = (cg[1:] < lo * (1 - 1e-12)) | (cg[:-1] > hi * (1 + 1e-12))
        assert f10[iz][outside].max() == 0.0 and f10[iz].max() > 0, iz


@pytest.mark.parametrize("h,a,beta,gam,ect", [(6.0, 0.9, 0.0, 2.0, 300.0), (3.0, 0.998, 0.2, 1.6, 60.0)])
def test_relxilllp_primary_only_closed_form(relxlp, h, a, beta, gam, ect):
    """refl_frac = 0: the observed cut-off power law (the source's Ecut = Ecut /
    g shifted back by g), normalised at the source, times f_inf_rest / 0.5
    g^Gamma (and the Doppler factor squared for a moving source): the integral
    over each caller bin, 1e-3."""
    T, xt = relxlp
    p = list(RELXILLLP_P)
    p[0], p[1], p[2], p[7], p[10], p[11] = h, beta, a, gam, ect, 0.0
    e = log_grid(200, 0.1, 100.0)
    y, _, st = oracle.relxill_model(oracle.MOD_RELXILLLP, T, xt, p, e)
    assert st == 0
    rp, _ = oracle.rel_params(oracle.MOD_RELXILLLP, p)
    sp, _ = oracle.system_parameters(T, rp)
    f_inf_rest = sp["refl"][4]
    g = math.sqrt(1 - 2 * h / (h * h + a * a))
    dop = 1.0
    if beta >= 1e-4:
        dop = math.sqrt(1 - beta * beta) / (1 - beta * math.cos(math.radians(p[3])))
        g *= dop
    nf = oracle.primary_norm_factor(gam, ect / g)
    fac = f_inf_rest / 0.5 * g ** gam * dop ** 2 * nf
    want = np.array([fac * quad(lambda x: math.exp(1 / ect) * x ** -gam * math.exp(-x / ect), a0, b0,
                                epsrel=1e-10)[0] for a0, b0 in zip(e[:-1], e[1:])])
    assert rel_err(y, want) < 1e-3


def test_relxilllp_refl_frac_and_boost(relxlp):
    """The reflection scales with |refl_frac| over the lamp post's reflection
    fraction; with switch_reflfrac_boost = 1, with |refl_frac| alone; the
    primary only for refl_frac >= 0."""
    T, xt = relxlp
    e = log_grid(300, 0.1, 100.0)

    def run(rf, boost=0.0):
        p = list(RELXILLLP_P)
        p[11], p[13] = rf, boost
        return oracle.relxill_model(oracle.MOD_RELXILLLP, T, xt, p, e)[0]
    r1, r2, p0, p3 = run(-1.0), run(-2.0), run(0.0), run(3.0)
    assert r1.max() > 0 and p0.max() > 0
    assert rel_err(r2, 2 * r1) < 1e-14
    assert rel_err(p3, 3 * r1 + p0) < 1e-13
    rp, _ = oracle.rel_params(oracle.MOD_RELXILLLP, RELXILLLP_P)
    sp, _ = oracle.system_parameters(T, rp)
    assert rel_err(run(-1.0, 1.0), sp["refl"][0] * r1) < 1e-13
