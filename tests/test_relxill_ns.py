"""The blackbody-primary relxill model in the oracle (relxillNS: relxill_kernel
with prim_type PRIM_SPEC_BB, src/ModelDefinition.cpp:89, :357-376;
spec_blackbody, src/Xillspec.cpp:255-268, normalised as the xillver
incident spectrum, :164-205, :302-332; the xillverNS table, src/xilltable.c:
742-748), pinned by closed forms: the blackbody's normalisation (its energy
flux over all energies is pi^4/15 whatever kT), the primary alone against
quadrature, the reflection against the relxill-style convolution of the
table's spectra at the named parameters (which pins the parameter order),
the kTbb axis on the table's nodes, and the refl_frac rule.  CPU only."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import oracle
import relxill_amd as rx
from tests.helpers import log_grid, rel_err, tables

KEV2ERG = 1.602177e-09
NORM_XILLVER = 1e15 / 4 / math.pi

# relxillNS: Index1 Index2 Rbr a Incl Rin Rout z kTbb logxi Afe logN refl_frac
NS_P = [3.4, 2.6, 12.0, 0.9, 40.0, -1.3, 350.0, 0.0, 1.7, 2.9, 2.2, 16.3, 2.0]


@pytest.fixture(scope="module")
def ns():
    _, _, T = tables()
    rx.xillnstable_synthetic(7, 6, 4, 5, 4, 10, 600)
    x = rx.xilltable_view(blackbody=True)
    return T, x, oracle.Xill(x)


def _bb_integrand(e, kt):
    return e ** 3 / (kt ** 4 * math.expm1(min(e / kt, 700.0)))


@pytest.mark.parametrize("kt", [0.5, 2.0, 7.0])
def test_blackbody_normalisation(kt):
    """1 / calcNormWrtXillverTableSpec of the blackbody: the 500-bin midpoint
    sum of E times its photons is the integral of E^3 / (kT^4 (e^(E/kT) - 1))
    over 0.1-1000 keV to 2e-4; that integral is pi^4/15 up to the part below
    0.1 keV (x^3 / (e^x - 1) integrates to pi^4/15 for every kT).  A dropped
    kT^4, keV->erg factor or 4 pi misses."""
    nf = oracle.prim_norm_factor(0.0, 300.0, prim=2, ktbb=kt)
    band = quad(_bb_integrand, 0.1, 1000.0, args=(kt,), limit=400, epsabs=0, epsrel=1e-12)[0]
    low = quad(_bb_integrand, 0.0, 0.1, args=(kt,), limit=400, epsabs=0, epsrel=1e-12)[0]
    assert band + low == pytest.approx(math.pi ** 4 / 15, rel=1e-9)
    want = 1.0 / (band * 1e20 * KEV2ERG / NORM_XILLVER)
    print(f"kT {kt}: nf / quadrature - 1 = {nf / want - 1:.2e}")
    assert nf / want - 1 == pytest.approx(0.0, abs=2e-4)


@pytest.mark.parametrize("kt,z", [(1.7, 0.0), (0.6, 0.4), (6.0, 1.5)])
def test_relxillns_primary_only(ns, kt, z):
    """refl_frac = 0: the normalised blackbody integrated over the redshifted
    caller bins; no energy shift of the blackbody.  The model takes midpoint
    values on the 4096-bin grid and splits partial bins linearly when it
    rebins, a first-order error of (dE/kT) / (2 x 9 grid bins per caller bin)
    on the Wien tail: 3e-3 over the bins above 1e-10 of the maximum.  A
    dropped kT^4, a shifted blackbody or another slot miss by far more."""
    T, _, xt = ns
    p = list(NS_P)
    p[7], p[8], p[12] = z, kt, 0.0
    e = log_grid(200, 0.1, 100.0)
    y, _, st = oracle.relxill_model(oracle.MOD_RELXILLNS, T, xt, p, e)
    assert st == 0
    nf = oracle.prim_norm_factor(0.0, 300.0, prim=2, ktbb=kt)
    es = e * (1 + z)
    want = np.array([nf * quad(lambda x: x * x / (kt ** 4 * math.expm1(min(x / kt, 700.0))), a, b,
                               epsabs=0, epsrel=1e-10)[0] for a, b in zip(es[:-1], es[1:])])
    err = rel_err(y, want)
    print(f"relxillNS primary (kT {kt}, z {z}): {err:.2e}")
    assert err < 3e-3


def test_relxillns_reflection_is_the_angle_resolved_convolution(ns):
    """refl_frac = -1: the relxill-style convolution of the table's per-angle
    spectra at (kTbb, A_Fe, logXi, logN) with relconv's geometry, 1e-13 and
    the same angle distribution; every parameter differs from the others, so
    a slot out of the lmodel.dat order fails."""
    T, _, xt = ns
    e = log_grid(300, 0.1, 100.0)
    for p in (NS_P[:12] + [-1.0], [2.0, 4.0, 8.0, -0.5, 70.0, 3.0, 300.0, 0.0, 3.3, 1.2, 5.1, 18.2, -1.0]):
        y, dist, st = oracle.relxill_model(oracle.MOD_RELXILLNS, T, xt, p, e)
        assert st == 0
        xp = oracle.xill_param(0.0, p[10], p[9], 300.0, dens=p[11], prim=2, ktbb=p[8])
        ex, flu = oracle.xill_spectra(xt, xp)
        cg = log_grid(4096, 0.00035, 2000.0)
        ref, dist_ref, st = oracle.relxill_conv(T, p[:7] + [0.0], cg, ex, flu, e)
        assert st == 0
        assert np.array_equal(dist.reshape(-1), dist_ref)
        assert rel_err(y, ref) < 1e-13


def test_kTbb_axis_on_the_nodes(ns):
    """On the table's nodes the interpolation returns the stored rows divided
    by 10^logXi and 10^(logN - 15) of the node (each division rounded to
    float), for every inclination: kTbb selects the first axis."""
    _, x, xt = ns
    val = x["val"]
    nv = x["n_val"]
    for it, ia, ix, idn in ((0, 0, 0, 0), (3, 2, 4, 1), (5, 3, 2, 3), (2, 1, 1, 2)):
        xp = oracle.xill_param(0.0, float(val[1][ia]), float(val[2][ix]), 300.0, dens=float(val[3][idn]),
                               prim=2, ktbb=float(val[0][it]))
        _, flu = oracle.xill_spectra(xt, xp)
        row0 = (((it * nv[1] + ia) * nv[2] + ix) * nv[3] + idn) * nv[4]
        for k in range(nv[4]):
            r = x["spec"][row0 + k].astype(np.float32)
            r = (r / np.float64(10.0 ** float(val[2][ix]))).astype(np.float32)
            if abs(float(val[3][idn]) - 15) > 1e-6:
                r = (r / np.float64(10.0 ** (float(val[3][idn]) - 15))).astype(np.float32)
            assert np.array_equal(flu[k], r.astype(np.float64)), (it, ia, ix, idn, k)


def test_relxillns_linear_in_refl_frac(ns):
    """The reflection scales with |refl_frac|; the primary only for
    refl_frac >= 0."""
    T, _, xt = ns
    e = log_grid(300, 0.1, 100.0)
    run = lambda rf: oracle.relxill_model(oracle.MOD_RELXILLNS, T, xt, NS_P[:12] + [rf], e)[0]
    r1, r2, p0, p3 = run(-1.0), run(-2.0), run(0.0), run(3.0)
    assert r1.max() > 0 and p0.max() > 0
    assert rel_err(r2, 2 * r1) < 1e-14
    assert rel_err(p3, 3 * r1 + p0) < 1e-13
