"""GPU edge cases against the oracle: grids that miss the line, one- and
two-bin grids, extreme spin / inclination / outer radius, redshift, broken
power laws with limb darkening, the largest caller grids, convolution inputs
that the reference rejects, and per-request statuses inside a batch."""
import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import RELLINE_CFG0, log_grid, rel_err, tables

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tabs():
    rel, lp, T = tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    return T


def both(tabs, model, p, e, fin=None):
    ref, st_ref = oracle.eval_model(model, tabs, p, e, flux_in=fin)
    out = {}
    for prec in (rx.FP64, rx.FP32):
        out[prec] = rx.eval_model(model, p, e, flux_in=fin, precision=prec, raise_on_error=False)
    return ref, st_ref, out


def test_grid_missing_the_line_is_nan_like_the_reference(tabs):
    """renorm_relline_profile divides by the profile sum (src/Relprofile.cpp:757-779);
    with no overlap that sum is 0 and the reference returns NaN."""
    ref, st, out = both(tabs, 1, RELLINE_CFG0, log_grid(100, 20.0, 30.0))
    assert st == 0 and np.isnan(ref).all()
    for prec, (y, s) in out.items():
        assert s == 0 and np.isnan(y).all(), prec


@pytest.mark.parametrize("edges", [[6.0, 7.0], [1.0, 6.4, 10.0], [6.39, 6.41]])
def test_tiny_grids(tabs, edges):
    e = np.array(edges, dtype=np.float64)
    ref, st, out = both(tabs, 1, RELLINE_CFG0, e)
    assert st == 0
    for prec, (y, s) in out.items():
        assert s == 0 and rel_err(y, ref) < (1e-8 if prec == rx.FP64 else 1e-4), (prec, y, ref)


@pytest.mark.parametrize("p", [
    [6.4, 3.0, 3.0, 15.0, -0.998, 87.0, -1.0, 1000.0, 0.0, 0.0],   # retrograde, edge-on, Rout at the table edge
    [6.4, 3.0, 3.0, 15.0, 0.998, 3.0, -1.0, 400.0, 0.5, 0.0],      # face-on, redshift 0.5
    [6.4, 5.0, 2.0, 6.0, 0.9, 60.0, -1.0, 400.0, 0.0, 2.0],        # broken power law, Haardt limb
    [1.0, -1.5, 4.5, 40.0, 0.3, 45.0, -3.0, -50.0, 0.0, 1.0],      # negative radii in ISCO units, Laor limb
])
def test_extreme_parameters(tabs, p):
    e = log_grid(2500, 0.05, 12.0)
    ref, st, out = both(tabs, 1, p, e)
    assert st == 0
    for prec, (y, s) in out.items():
        assert s == 0
        assert rel_err(y, ref) < (1e-8 if prec == rx.FP64 else 1e-4), (prec, rel_err(y, ref))


def test_largest_caller_grid(tabs):
    e = log_grid(8000, 0.1, 10.0)
    ref, st, out = both(tabs, 1, RELLINE_CFG0, e)
    for prec, (y, s) in out.items():
        assert s == 0 and rel_err(y, ref) < (1e-8 if prec == rx.FP64 else 1e-4)


def test_relconv_rejects_empty_input_and_limb_conv_matches(tabs):
    e = log_grid(500, 0.1, 100.0)
    y, st = rx.eval_model(11, [3, 3, 15, 0.998, 30, -1, 400, 0], e, flux_in=np.zeros(500), raise_on_error=False)
    assert st == rx.RX_ERR_INPUT and np.all(y == 0)   # src/LocalModel.cpp:84-86
    fin = np.exp(-np.linspace(0, 5, 500))
    ref, st, out = both(tabs, 11, [3, 3, 15, 0.998, 30, -1, 400, 1], e, fin)
    assert st == 0
    for prec, (y, s) in out.items():
        assert s == 0 and rel_err(y, ref) < (1e-8 if prec == rx.FP64 else 1e-4)


def test_batch_statuses_are_per_request(tabs):
    e = log_grid(600, 0.1, 10.0)
    bad = list(RELLINE_CFG0)
    bad[5] = 89.0  # inclination outside 3-87 deg
    reqs = [dict(model=1, energy=e, params=RELLINE_CFG0), dict(model=1, energy=e, params=bad),
            dict(model=11, energy=e, params=[3, 3, 15, 0.5, 30, -1, 400, 0], flux_in=np.zeros(600)),
            dict(model=1, energy=log_grid(50, 20.0, 30.0), params=RELLINE_CFG0)]
    b = rx.Batch(reqs)
    b.run(rx.FP32)
    b.sync()
    ref0, _ = oracle.eval_model(1, tabs, RELLINE_CFG0, e)
    y0, s0 = b.output(0)
    y1, s1 = b.output(1)
    y2, s2 = b.output(2)
    y3, s3 = b.output(3)
    assert s0 == 0 and rel_err(y0, ref0) < 1e-4
    assert s1 == rx.RX_ERR_PARAM and np.all(y1 == 0)
    assert s2 == rx.RX_ERR_INPUT and np.all(y2 == 0)
    assert s3 == 0 and np.isnan(y3).all()
    b.close()


def test_stale_batch_refuses_to_run(tabs):
    """A batch's device buffers are sized for the configuration it was built
    for; after a change of the fine-grid size it reports an error instead of
    running, and a rebuilt batch runs."""
    e = log_grid(600, 0.1, 10.0)
    b = rx.Batch([dict(model=1, energy=e, params=RELLINE_CFG0)])
    b.run(rx.FP32)
    b.sync()
    assert rx.lib.rx_set_nr_fine(500) == 0
    try:
        with pytest.raises(rx.RxError):
            b.run(rx.FP32)
    finally:
        assert rx.lib.rx_set_nr_fine(1000) == 0
    b.close()
    b = rx.Batch([dict(model=1, energy=e, params=RELLINE_CFG0)])
    b.run(rx.FP32)
    b.sync()
    ref, _ = oracle.eval_model(1, tabs, RELLINE_CFG0, e)
    y, st = b.output(0)
    assert st == 0 and rel_err(y, ref) < 1e-4
    b.close()
