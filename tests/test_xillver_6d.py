"""Six-parameter xillver tables in the oracle (new_xillTable accepts five or
six parameters, src/xilltable.c:70-81; the first is interpolated outside the
other five, interp_6d_tab_incl / interp_6d_tab, :997-1045, with the rows'
density division of normalizeXillverSpecLogxiDensity, :488-511), pinned by
closed forms: an affine table is reproduced in all six parameters, a first
parameter on a node gives the five-parameter interpolation of that node's
slice bit for bit (and the lerp of two slices between nodes), and the
relxill model on the table at a node equals the model on the slice.  The
library's six-parameter synthetic table and its view are checked too.  CPU
only."""
import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import AFFINE_XILL_C, affine_xilltable6, log_grid, rel_err, tables, xill_slice

C_DENS = 0.053


@pytest.fixture(scope="module")
def affine6():
    x, g = affine_xilltable6(c_dens=C_DENS)
    return x, g, oracle.Xill(x)


def _expected(x, g, gam, afe, lxi, ect, dens, kk=None):
    v = x["val"]
    cl = lambda q, k: min(max(q, float(v[k][0])), float(v[k][-1]))
    c = AFFINE_XILL_C
    k = np.arange(x["n_val"][5]) if kk is None else kk
    lin = (c[0] + c[1] * cl(gam, 0) + c[2] * cl(afe, 1) + c[3] * cl(lxi, 2) + c[4] * cl(ect, 3) + c[5] * k
           + C_DENS * cl(dens, 4))
    return np.multiply.outer(lin, g)


@pytest.mark.parametrize("dens", [15.0, 15.4, 16.9, 18.37, 19.0, 14.0, 20.0])
def test_affine_in_all_six_parameters(affine6, dens):
    """orc_xill_spectra on a table affine in (Gamma, A_Fe, logXi, Ecut, Dens)
    and the inclination index returns the affine value at every inclination
    (the parameters clamped into the table): 3e-6 (float rows, two
    float-rounded divisions)."""
    x, g, xt = affine6
    for gam, afe, lxi, ect in ((2.1, 1.3, 3.1, 250.0), (1.6, 6.0, 0.5, 30.0), (3.5, 0.4, 4.7, 3.0)):
        _, flu = oracle.xill_spectra(xt, oracle.xill_param(gam, afe, lxi, ect, dens=dens))
        want = _expected(x, g, gam, afe, lxi, ect, dens)
        assert np.abs(flu / want - 1).max() < 3e-6, (dens, gam)


def test_affine_with_the_inclination_interpolated(affine6):
    """orc_xill_spectrum_incl (the xillver models: the inclination an inner
    factor too, interp_6d_tab) on the same table: the affine value at the
    inclination's fractional index, 3e-6."""
    x, g, xt = affine6
    v = np.asarray(x["val"][5], dtype=np.float32)
    for incl, gam, dens in ((33.3, 2.2, 16.2), (61.2, 1.3, 18.9), (5.0, 3.0, 15.0)):
        q = np.float32(min(max(np.float32(incl), v[0]), v[-1]))
        i = int(min(max(np.searchsorted(v, q, side="right") - 1, 0), v.size - 2))
        f = float(np.float32((q - v[i]) / (v[i + 1] - v[i])))
        _, flu = oracle.xill_spectrum_incl(xt, oracle.xill_param(gam, 2.0, 2.5, 100.0, dens=dens, incl=incl))
        want = _expected(x, g, gam, 2.0, 2.5, 100.0, dens, kk=i + f)
        assert np.abs(flu / want - 1).max() < 3e-6, (incl, dens)


@pytest.fixture(scope="module")
def synth6():
    rx.xillcptable_synthetic_dens(8, 5, 6, 4, 5, 5, 8, 400)
    try:
        x6 = rx.xilltable_view(nthcomp=True)
    finally:
        rx.xillcptable_synthetic()
    return x6


def test_library_six_parameter_table(synth6):
    """rx_xillcptable_synthetic_dens: Gamma, A_Fe, logXi, kTe, Dens (15-19),
    Incl; rows as stored: the block of logN = 18 is that of 15 times 10^3
    (1 + 0.3 exp(-E / 1.5 keV)), the file scale and the soft excess (float
    rows: 1e-6)."""
    x6 = synth6
    assert x6["n_par"] == 6 and x6["pid"] == [0, 1, 2, 3, 4, 7]
    assert x6["n_val"] == [6, 4, 5, 5, 5, 8]
    assert np.allclose(x6["val"][4], [15, 16, 17, 18, 19])
    spec = x6["spec"].reshape(6, 4, 5, 5, 5, 8, 400).astype(np.float64)
    em = 0.5 * (x6["elo"].astype(np.float64) + x6["ehi"])
    band = (em > 1.0) & (em < 20.0)
    lo, hi = spec[..., 0, :, :][..., band], spec[..., 3, :, :][..., band]
    ok = lo > 0
    want = 1e3 * (1.0 + 0.3 * np.exp(-em[band] / 1.5))
    assert ok.mean() > 0.9
    assert np.abs(hi[ok] / lo[ok] / np.broadcast_to(want, lo.shape)[ok] - 1).max() < 1e-6


# kTe inside 15-19: neither the six-parameter table (whose axis 3 is kTe) nor
# its five-parameter slices (whose axis 3 is Dens) snap the kTe factor in
# ensure_ecut_within_boundarys (src/xilltable.c:1054-1066), which reads the
# table's axis 3 whatever it holds
KTE = 17.0


def test_first_axis_node_equals_the_slice(synth6):
    """Gamma on a node of the first axis: the six-parameter interpolation is
    1 x (the five-parameter interpolation of that node's slice) + 0 x (the
    next one), bit for bit the slice's, for every inclination; between two
    nodes it is f s2 + (1 - f) s1 of the two slices (f the float factor of
    the bracket).  A mixed-up outer index, stride or factor fails."""
    x6 = synth6
    xt6 = oracle.Xill(x6)
    p = dict(afe=1.7, lxi=2.9, ect=KTE, dens=16.4, prim=1)
    for i in range(6):
        g0 = float(x6["val"][0][i])
        xs = oracle.Xill(xill_slice(x6, i))
        _, f6 = oracle.xill_spectra(xt6, oracle.xill_param(g0, **p))
        _, f5 = oracle.xill_spectra(xs, oracle.xill_param(g0, **p))
        assert np.array_equal(f6, f5), i
        if i < 5:
            g1 = float(x6["val"][0][i + 1])
            gm = 0.3 * g0 + 0.7 * g1
            xp = oracle.xill_param(gm, **p)
            ind, fac = oracle.xill_corner(xt6, xp)
            assert ind[0] == i and 0.6 < fac[0] < 0.8
            _, fm = oracle.xill_spectra(xt6, xp)
            _, fa = oracle.xill_spectra(xs, oracle.xill_param(g0, **p))
            _, fb = oracle.xill_spectra(oracle.Xill(xill_slice(x6, i + 1)), oracle.xill_param(g1, **p))
            assert np.array_equal(fm, fac[0] * fb + (1.0 - fac[0]) * fa), i


def test_relxillcp_on_the_slice(synth6):
    """relxillCp with Gamma on a node of the six-parameter table's first axis
    is relxillCp on that node's five-parameter slice (the same rows, factors
    and density division; the primary does not read the table): 1e-15."""
    _, _, T = tables()
    x6 = synth6
    xt6 = oracle.Xill(x6)
    e = log_grid(300, 0.1, 100.0)
    for i, logn in ((2, 17.0), (4, 15.6)):
        g0 = float(x6["val"][0][i])
        p = [40.0, 0.7, -1.0, 400.0, 15.0, 3.0, 2.5, 0.0, g0, 2.7, logn, 3.0, KTE, 1.5]
        y6, _, st = oracle.relxill_model(oracle.MOD_RELXILLCP, T, xt6, p, e)
        assert st == 0
        y5, _, st = oracle.relxill_model(oracle.MOD_RELXILLCP, T, oracle.Xill(xill_slice(x6, i)), p, e)
        assert st == 0
        assert rel_err(y6, y5) < 1e-15, i
