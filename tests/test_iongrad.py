"""relxilllpCp's power-law ionisation gradient in the oracle
(IonGradient::calculate_gradient, src/IonGradient.cpp:340-395: calc_ion_grad_pl,
:174-182, then adjust_parameter_to_be_within_xillver_bounds, :36-44, :388, on
N_ZONES_IONGRAD = 25 zones, src/relutility.c:517-527, src/common.h:135),
pinned by closed forms: every zone's logXi is log(exp(logXi0) (r_z / r_0)^-index)
of its mean radius, kept within 0-4.7; the zones cover Rin-Rout; and on a
table whose spectra do not depend on logXi the gradient changes nothing.
CPU only."""
import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import log_grid, rel_err, tables

# relxilllpCp: Incl a Rin Rout h beta gamma logxi logN Afe kTe refl_frac z
#              iongrad_index iongrad_type switch_returnrad switch_reflfrac_boost
LPCP_G = [33.0, 0.9, -1.0, 400.0, 6.0, 0.0, 2.0, 3.1, 15.0, 1.0, 60.0, 1.0, 0.0, 0.7, 1.0, 0.0, 0.0]


@pytest.fixture(scope="module")
def cp():
    _, _, T = tables()
    rx.xillcptable_synthetic(6, 5, 4, 6, 5, 10, 600)
    return T, rx.xilltable_view(nthcomp=True)


@pytest.mark.parametrize("lxi0,index", [(3.1, 0.7), (2.0, 0.0), (4.7, -0.5), (0.3, 1.5), (4.2, 3.0)])
def test_zone_logxi_power_law(cp, lxi0, index):
    """25 zones from Rin to Rout; zone z's logXi is
    log(exp(logXi0) (rmean_z / rmean_0)^-index) clamped into 0-4.7 (the
    natural log of the reference's expression, 1e-14); index 0 leaves every
    zone at logXi0."""
    T, _ = cp
    p = list(LPCP_G)
    p[7], p[13] = lxi0, index
    z, st = oracle.relxill_zones(oracle.MOD_RELXILLLPCP, T, p)
    assert st == 0 and z.shape == (25, 7)
    assert np.array_equal(z[1:, 0], z[:-1, 1]) and np.all(np.diff(z[:, 0]) > 0)
    assert z[0, 0] == pytest.approx(float(oracle.lib.orc_kerr_rms(p[1])), rel=1e-14)
    assert z[-1, 1] == pytest.approx(400.0, rel=1e-14)
    rm = 0.5 * (z[:, 0] + z[:, 1])
    want = np.clip(np.log(np.exp(lxi0) * (rm / rm[0]) ** (-index)), 0.0, 4.7)
    assert np.abs(z[:, 6] - want).max() < 1e-14
    if index == 0.0:
        assert np.all(z[:, 6] == lxi0)
    if (lxi0, index) == (4.7, -0.5):
        assert np.all(z[1:, 6] == 4.7)
    if (lxi0, index) == (0.3, 1.5):
        assert abs(z[0, 6] - 0.3) < 1e-15 and z[-1, 6] == 0.0


def test_gradient_enters_through_logxi_only(cp):
    """On a table whose spectra, once divided by 10^logXi, are the same at
    every logXi node, the gradient changes the model by no more than the
    float rounding of the rows (1e-6); on the table itself it changes the
    reflection by more than 1%.  The angle distribution holds 25 zones, each summing to 1."""
    T, x = cp
    e = log_grid(200, 0.1, 100.0)
    nv = x["n_val"]
    spec = x["spec"].reshape(*nv, x["n_ener"]).astype(np.float64)
    lxi = x["val"][2].astype(np.float64)
    flat = spec.copy()
    for ix in range(nv[2]):
        flat[:, :, ix] = spec[:, :, 0] * 10.0 ** (lxi[ix] - lxi[0])
    xf = dict(x)
    xf["spec"] = flat.reshape(-1, x["n_ener"]).astype(np.float32)
    xt, xtf = oracle.Xill(x), oracle.Xill(xf)
    p0, p1 = list(LPCP_G), list(LPCP_G)
    p0[13], p1[13] = 0.0, 1.3
    p0[11] = p1[11] = -1.0  # the reflection alone
    y0, d0, st0 = oracle.relxill_model(oracle.MOD_RELXILLLPCP, T, xtf, p0, e)
    y1, d1, st1 = oracle.relxill_model(oracle.MOD_RELXILLLPCP, T, xtf, p1, e)
    assert st0 == 0 and st1 == 0
    assert rel_err(y1, y0) < 1e-6
    assert d1.shape == (25, nv[-1]) and np.allclose(d1.sum(axis=1), 1.0, rtol=1e-12)
    g0, _, _ = oracle.relxill_model(oracle.MOD_RELXILLLPCP, T, xt, p0, e)
    g1, _, _ = oracle.relxill_model(oracle.MOD_RELXILLLPCP, T, xt, p1, e)
    assert rel_err(g1, g0) > 1e-2
