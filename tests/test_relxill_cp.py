"""The nthcomp-primary relxill models in the oracle (relxillCp, relxilllpCp:
relxill_kernel with prim_type PRIM_SPEC_NTHCOMP, src/ModelDefinition.cpp:
38-50, :88, :357-368; src/Xillspec.cpp:243-332; src/Relxill.cpp:197-200),
pinned without the reference (it needs cfitsio): the primary alone against
nthcomp and an independent restatement of its normalisation, the reflection
of relxillCp against relxill's on the same table (which pins the parameter
order), the lamp post's kTe frame and zones, the density division of a table
without a density axis, and the unsupported ionisation gradients.  CPU only."""
import math

import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import log_grid, rel_err, tables

KEV2ERG = 1.602177e-09
NORM_XILLVER = 1e15 / 4 / math.pi

# relxillCp: Incl a Rin Rout Rbr Index1 Index2 z gamma logxi logN Afe kTe refl_frac
CP_P = [37.0, 0.93, -1.7, 310.0, 12.0, 3.6, 2.4, 0.0, 1.85, 2.7, 15.0, 2.3, 45.0, 3.0]
# relxilllpCp: Incl a Rin Rout h beta gamma logxi logN Afe kTe refl_frac z
#              iongrad_index iongrad_type switch_returnrad switch_reflfrac_boost
LPCP_P = [33.0, 0.9, -1.0, 400.0, 6.0, 0.0, 2.0, 3.1, 15.0, 1.0, 60.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]


@pytest.fixture(scope="module")
def cp():
    _, _, T = tables()
    rx.xillcptable_synthetic(6, 5, 4, 6, 5, 10, 600)
    return T, oracle.Xill(rx.xilltable_view(nthcomp=True))


def _fine(e, kte, gam, shift=1.0, sub=16):
    """nthcomp photons per bin of e from 16 sub-bins per bin: its trapezoid
    error on steep tails falls below 1e-5"""
    le = np.log(e)
    fine = np.exp(np.concatenate([np.linspace(le[i], le[i + 1], sub + 1)[:-1] for i in range(e.size - 1)] + [le[-1:]]))
    fine[::sub] = e
    return oracle.nthcomp(fine, kte, gam, shift).reshape(-1, sub).sum(axis=1)


def _nf_nthcomp(kte, gam):
    """1 / calcNormWrtXillverTableSpec of nthcomp on 500 log bins over
    0.1-1000 keV, restated here: sum of photons x bin centre x keV->erg x 1e20
    over 1e15 / (4 pi)"""
    g = log_grid(500, 0.1, 1000.0)
    f = oracle.nthcomp(g, kte, gam)
    return 1.0 / (np.sum(f * 0.5 * (g[1:] + g[:-1]) * 1e20 * KEV2ERG) / NORM_XILLVER)


@pytest.mark.parametrize("gam,kte,z", [(2.0, 60.0, 0.0), (1.6, 20.0, 0.3), (2.6, 250.0, 1.2)])
def test_relxillcp_primary_only(cp, gam, kte, z):
    """refl_frac = 0: nthcomp (kTe, Gamma) normalised as the xillver incident
    spectrum, on the redshifted caller grid.  The model evaluates nthcomp on
    the 4096-bin convolution grid and rebins; integrated on 16 sub-bins of
    each caller bin it agrees to 2e-4.  Using Ecut's cut-off power law, kTe
    from another slot or the cut-off power law's norm misses by far more."""
    T, xt = cp
    p = list(CP_P)
    p[7], p[8], p[12], p[13] = z, gam, kte, 0.0
    e = log_grid(200, 0.1, 100.0)
    y, _, st = oracle.relxill_model(oracle.MOD_RELXILLCP, T, xt, p, e)
    assert st == 0
    want = _fine(e * (1 + z), kte, gam) * _nf_nthcomp(kte, gam)
    err = rel_err(y, want)
    print(f"relxillCp primary (Gamma {gam}, kTe {kte}, z {z}): {err:.2e}")
    assert err < 2e-4


@pytest.mark.parametrize("h,a,beta,gam,kte", [(6.0, 0.9, 0.0, 2.0, 60.0), (3.0, 0.998, 0.2, 1.6, 150.0)])
def test_relxilllpcp_primary_only(cp, h, a, beta, gam, kte):
    """refl_frac = 0: nthcomp with the source's kTe (not divided by g as
    relxilllp's Ecut) observed at redshift 1/g - 1, normalised at the source,
    times f_inf_rest / 0.5 g^Gamma (and the Doppler factor squared for a
    moving source), 2e-4."""
    T, xt = cp
    p = list(LPCP_P)
    p[4], p[5], p[1], p[6], p[10], p[11] = h, beta, a, gam, kte, 0.0
    e = log_grid(200, 0.1, 100.0)
    y, _, st = oracle.relxill_model(oracle.MOD_RELXILLLPCP, T, xt, p, e)
    assert st == 0
    # the same geometry through relxilllp's parameters
    lp = [h, beta, a, p[0], p[2], p[3], 0.0, gam, 3.1, 1.0, 300.0, 0.0, 0.0, 0.0]
    rp, _ = oracle.rel_params(oracle.MOD_RELXILLLP, lp)
    sp, _ = oracle.system_parameters(T, rp)
    f_inf_rest = sp["refl"][4]
    g = math.sqrt(1 - 2 * h / (h * h + a * a))
    dop = 1.0
    if beta >= 1e-4:
        dop = math.sqrt(1 - beta * beta) / (1 - beta * math.cos(math.radians(p[0])))
        g *= dop
    fac = f_inf_rest / 0.5 * g ** gam * dop ** 2 * _nf_nthcomp(kte, gam)
    want = _fine(e, kte, gam, shift=g) * fac
    err = rel_err(y, want)
    print(f"relxilllpCp primary (h {h}, a {a}, beta {beta}): {err:.2e}")
    assert err < 2e-4


def test_relxillcp_reflection_is_relxills_on_the_same_table(cp):
    """refl_frac < 0 (reflection only) for the broken power law: no primary,
    and the normalisation change is 1 (no energy shift), so relxillCp is
    relxill on the same table with kTe as Ecut, bit for bit.  Every parameter
    differs from the others, so a slot out of the lmodel.dat order shows."""
    T, xt = cp
    e = log_grid(300, 0.1, 100.0)
    for rf in (-1.0, -2.5):
        p = list(CP_P)
        p[13] = rf
        y, d, st = oracle.relxill_model(oracle.MOD_RELXILLCP, T, xt, p, e)
        # relxill: Index1 Index2 Rbr a Incl Rin Rout z gamma logxi Afe Ecut refl_frac
        q = [p[5], p[6], p[4], p[1], p[0], p[2], p[3], p[7], p[8], p[9], p[11], p[12], p[13]]
        y2, d2, st2 = oracle.relxill_model(oracle.MOD_RELXILL, T, xt, q, e)
        assert st == 0 and st2 == 0 and y.max() > 0
        assert np.array_equal(y, y2) and np.array_equal(d, d2)
    # the primary's absence: refl_frac = 0 gives the primary alone, a
    # positive refl_frac the sum
    p = list(CP_P)
    p[13] = 0.0
    prim, _, _ = oracle.relxill_model(oracle.MOD_RELXILLCP, T, xt, p, e)
    p[13] = -2.0
    refl, _, _ = oracle.relxill_model(oracle.MOD_RELXILLCP, T, xt, p, e)
    p[13] = 2.0
    both, _, _ = oracle.relxill_model(oracle.MOD_RELXILLCP, T, xt, p, e)
    assert rel_err(both, prim + refl) < 1e-12


def test_relxilllpcp_zones_and_kte_frame(cp):
    """The zones of relxilllpCp are relxilllp's (same geometry through the
    remapped slots); each sees the source's kTe times its energy shift (no
    division by g), and its normalisation change is the nthcomp one
    nf(kTe g_z) / nf(kTe)."""
    T, _ = cp
    p = list(LPCP_P)
    p[1], p[4], p[5], p[10] = 0.95, 4.0, 0.1, 80.0
    zc, st = oracle.relxill_zones(oracle.MOD_RELXILLLPCP, T, p)
    assert st == 0
    lp = [p[4], p[5], p[1], p[0], p[2], p[3], p[12], p[6], p[7], p[9], 300.0, p[11], p[15], p[16]]
    zl, st = oracle.relxill_zones(oracle.MOD_RELXILLLP, T, lp)
    assert st == 0
    assert np.array_equal(zc[:, :4], zl[:, :4])
    assert np.array_equal(zc[:, 4], 80.0 * zc[:, 3])
    nf0 = _nf_nthcomp(80.0, p[6])
    for iz in range(zc.shape[0]):
        assert zc[iz, 5] == pytest.approx(_nf_nthcomp(80.0 * zc[iz, 3], p[6]) / nf0, rel=1e-12)
    # the inner zones see a lower kTe: their change differs from 1
    assert abs(zc[0, 5] - 1) > 1e-3


def test_density_division_without_density_axis(cp):
    """A five-parameter table has no density axis: its rows are divided by
    10^(logN - 15) when loaded (normalizeXillverSpecLogxiDensity with the
    request's logN, src/xilltable.c:488-511, renorm_xill_spec :478-486), so
    the reflection scales by that factor up to the float rounding of the
    rows."""
    T, xt = cp
    e = log_grid(300, 0.1, 100.0)
    p = list(CP_P)
    p[13] = -1.0
    y15, _, _ = oracle.relxill_model(oracle.MOD_RELXILLCP, T, xt, p, e)
    p[10] = 16.3
    y16, _, st = oracle.relxill_model(oracle.MOD_RELXILLCP, T, xt, p, e)
    assert st == 0
    assert rel_err(y16 * 10 ** 1.3, y15) < 1e-6


def test_alpha_ionisation_gradient_is_unsupported(cp):
    """iongrad_type 2 asks for the alpha-disc ionisation gradient, which is
    not built: status 2 (RX_ERR_UNSUPPORTED) and zero flux."""
    T, xt = cp
    p = list(LPCP_P)
    p[14] = 2.0
    y, _, st = oracle.relxill_model(oracle.MOD_RELXILLLPCP, T, xt, p, log_grid(50, 0.1, 100.0))
    assert st == 2 and not y.any()
