"""The non-relativistic xillver models in the oracle (xillver, xillverCp,
xillverNS: LocalModel::xillver_model, src/LocalModel.cpp:104-128, with
interp_5d_tab, src/xilltable.c:878-995, norm_xillver_spec,
src/Xillspec.cpp:536-545, and add_primary_component, src/Relbase.cpp:294-344),
pinned by closed forms on an affine table: the fifth interpolation in the
inclination, the clamp of the inclination, the table spectrum times
0.5 cos(Incl) on a caller grid made of table edges, the primary alone
against quadrature, the parameter orders and the refl_frac rule.  CPU only."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import oracle
import relxill_amd as rx
from tests.helpers import AFFINE_XILL_C, affine_xilltable, log_grid, rel_err


@pytest.fixture(scope="module")
def affine():
    x, g = affine_xilltable()
    return x, g, oracle.Xill(x)


def _incl_factor(v, incl):
    """the bracket and float factor of the inclination, clamped into the table"""
    v = np.asarray(v, dtype=np.float32)
    q = np.float32(min(max(np.float32(incl), v[0]), v[-1]))
    i = int(min(max(np.searchsorted(v, q, side="right") - 1, 0), v.size - 2))
    return i, float(np.float32((q - v[i]) / (v[i + 1] - v[i])))


def _affine(x, gam, afe, lxi, ect, kk, c=AFFINE_XILL_C):
    v = x["val"]
    cl = lambda q, k: min(max(q, float(v[k][0])), float(v[k][-1]))
    return c[0] + c[1] * cl(gam, 0) + c[2] * cl(afe, 1) + c[3] * cl(lxi, 2) + c[4] * cl(ect, 3) + c[5] * kk


@pytest.mark.parametrize("incl", [5.0, 33.3, 47.0, 61.2, 88.0, 2.0])
def test_inclination_is_the_fifth_interpolated_parameter(affine, incl):
    """On a table affine in (Gamma, A_Fe, logXi, Ecut) and in the inclination
    index k, interp_5d_tab gives the affine value at k = i + f, with (i, f)
    the inclination's bracket and float factor (clamped into the table at 2
    and 88 degrees): 2e-6."""
    x, g, xt = affine
    i, f = _incl_factor(x["val"][4], incl)
    for gam, afe, lxi, ect in ((2.1, 1.3, 3.1, 250.0), (1.6, 6.0, 0.5, 30.0)):
        _, flu = oracle.xill_spectrum_incl(xt, oracle.xill_param(gam, afe, lxi, ect, incl=incl))
        want = _affine(x, gam, afe, lxi, ect, i + f) * g
        assert np.abs(flu / want - 1).max() < 2e-6, (incl, gam)


def test_xillver_on_table_edges(affine):
    """xillver with refl_frac < 0 on a caller grid made of table edges (no
    redshift): every bin is the affine table value times 0.5 cos(Incl) times
    |refl_frac|, 2e-6 (the float table); with a redshift, the same spectrum
    on the grid divided by 1 + z, 1e-12."""
    x, g, xt = affine
    e = x["elo"][20:221].astype(np.float64)
    incl = 41.0
    i, f = _incl_factor(x["val"][4], incl)
    p = [2.3, 2.0, 120.0, 2.2, 0.0, incl, -2.5]
    y, st = oracle.xillver_model(oracle.MOD_XILLVER, xt, p, e)
    assert st == 0
    want = _affine(x, 2.3, 2.0, 2.2, 120.0, i + f) * g[20:220] * 0.5 * math.cos(math.radians(incl)) * 2.5
    assert rel_err(y, want) < 2e-6
    # z: the same spectrum on the grid divided by 1 + z
    z = 0.25
    p[4] = z
    y2, _ = oracle.xillver_model(oracle.MOD_XILLVER, xt, p, e / (1 + z))
    assert rel_err(y2, y) < 1e-12


@pytest.mark.parametrize("gam,ect,z", [(2.0, 300.0, 0.0), (1.5, 40.0, 0.6)])
def test_xillver_primary_only(affine, gam, ect, z):
    """refl_frac = 0: the normalised cut-off power law evaluated on the
    redshifted caller grid (midpoint per bin, as spec_cutoffpl), against
    quadrature: the midpoint rule's (d ln E)^2 / 24 terms on 1000 bins keep
    it within 2e-4; no reflection."""
    _, _, xt = affine
    e = log_grid(1000, 0.1, 100.0)
    y, st = oracle.xillver_model(oracle.MOD_XILLVER, xt, [gam, 1.0, ect, 3.1, z, 30.0, 0.0], e)
    assert st == 0
    nf = oracle.primary_norm_factor(gam, ect)
    es = e * (1 + z)
    want = np.array([nf * quad(lambda q: math.exp(1 / ect) * q ** -gam * math.exp(-q / ect), a, b,
                               epsrel=1e-10)[0] for a, b in zip(es[:-1], es[1:])])
    assert rel_err(y, want) < 2e-4


def test_parameter_orders_and_primaries(affine):
    """xillverCp and xillverNS read their lmodel.dat orders: with
    refl_frac < 0 (and the affine table, which has no kTe or kTbb axis and
    an Ecut axis) xillverCp's kTe takes Ecut's place, xillverNS keeps Ecut at
    300; with refl_frac = 0 each adds its own primary (nthcomp with its
    norm, the blackbody with its norm)."""
    x, g, xt = affine
    e = log_grid(150, 0.2, 80.0)
    ref, _ = oracle.xillver_model(oracle.MOD_XILLVER, xt, [2.2, 1.7, 55.0, 2.9, 0.1, 37.0, -1.3], e)
    cp, _ = oracle.xillver_model(oracle.MOD_XILLVERCP, xt, [2.2, 1.7, 55.0, 2.9, 15.0, 0.1, 37.0, -1.3], e)
    assert rel_err(cp, ref) < 1e-15
    ref300, _ = oracle.xillver_model(oracle.MOD_XILLVER, xt, [0.0, 1.7, 300.0, 2.9, 0.1, 37.0, -1.3], e)
    ns, _ = oracle.xillver_model(oracle.MOD_XILLVERNS, xt, [2.0, 1.7, 15.0, 2.9, 0.1, 37.0, -1.3], e)
    assert rel_err(ns, ref300) < 1e-15
    cp0, _ = oracle.xillver_model(oracle.MOD_XILLVERCP, xt, [2.2, 1.7, 55.0, 2.9, 15.0, 0.1, 37.0, 0.0], e)
    want = oracle.nthcomp(e * 1.1, 55.0, 2.2) * oracle.prim_norm_factor(2.2, 55.0, prim=1)
    assert rel_err(cp0, want) < 1e-14
    ns0, _ = oracle.xillver_model(oracle.MOD_XILLVERNS, xt, [2.0, 1.7, 15.0, 2.9, 0.1, 37.0, 0.0], e)
    want = oracle.blackbody(e * 1.1, 2.0) * oracle.prim_norm_factor(0.0, 300.0, prim=2, ktbb=2.0)
    assert rel_err(ns0, want) < 1e-14


def test_xillver_linear_in_refl_frac(affine):
    """The reflection scales with |refl_frac|; the primary only for
    refl_frac >= 0."""
    _, _, xt = affine
    e = log_grid(150, 0.2, 80.0)
    run = lambda rf: oracle.xillver_model(oracle.MOD_XILLVER, xt, [1.9, 2.0, 150.0, 3.0, 0.0, 50.0, rf], e)[0]
    r1, r2, p0, p3 = run(-1.0), run(-2.0), run(0.0), run(3.0)
    assert r1.max() > 0 and p0.max() > 0
    assert rel_err(r2, 2 * r1) < 1e-14
    assert rel_err(p3, 3 * r1 + p0) < 1e-13


def test_library_tables_carry_the_xillver_axes():
    """The three synthetic tables the library holds have the axes the models
    read (Gamma/A_Fe/logXi/Ecut, Gamma/A_Fe/logXi/kTe, kTbb/A_Fe/logXi/Dens,
    each with the inclination last)."""
    rx.xilltable_synthetic(5, 5, 4, 6, 5, 10, 300)
    rx.xillcptable_synthetic(6, 5, 4, 6, 5, 10, 300)
    rx.xillnstable_synthetic(7, 5, 4, 6, 5, 10, 300)
    try:
        assert rx.xilltable_view()["pid"] == [0, 1, 2, 3, 7]
        assert rx.xilltable_view(nthcomp=True)["pid"] == [0, 1, 2, 3, 7]
        assert rx.xilltable_view(blackbody=True)["pid"] == [5, 1, 2, 4, 7]
    finally:
        rx.xilltable_synthetic()
        rx.xillcptable_synthetic()
        rx.xillnstable_synthetic()
