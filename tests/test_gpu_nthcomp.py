"""The nthcomp primary continuum on the GPU (rx_nthcomp) against the FP64
oracle (spec_nthcomp, src/Xillspec.cpp:243-253; c_donthcomp,
src/donthcomp.c:683-793), and the GPU twins of the oracle's pins: the 1 keV
normalisation at every redshift and the redshift as a grid scaling."""
import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import log_grid

pytestmark = pytest.mark.gpu

CASES = [(60.0, 2.0, 1.0), (1.5, 1.4, 1.0), (400.0, 3.2, 0.8), (20.0, 1.9, 0.55), (150.0, 2.6, 1.3),
         (5.0, 1.7, 0.95), (1000.0, 2.2, 1.0), (0.3, 2.4, 1.0)]


@pytest.fixture(scope="module")
def dev():
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()


def _rel(a, b):
    scale = np.max(np.abs(b))
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12 * scale))


@pytest.mark.parametrize("grid", [(1000, 0.1, 1000.0), (4096, 0.00035, 2000.0), (37, 0.5, 50.0)])
def test_nthcomp_against_oracle(dev, grid):
    e = log_grid(*grid)
    kte = np.array([c[0] for c in CASES])
    gam = np.array([c[1] for c in CASES])
    sh = np.array([c[2] for c in CASES])
    y = rx.nthcomp(e, kte, gam, sh)
    worst = 0.0
    for k, (t, g, s) in enumerate(CASES):
        ref = oracle.nthcomp(e, t, g, s)
        assert np.all(np.isfinite(y[k]))
        err = _rel(y[k], ref)
        worst = max(worst, err)
        assert err < 1e-10, (t, g, s, err)
    print(f"grid {grid}: {len(CASES)} spectra, worst relative error per bin {worst:.2e}", flush=True)


def test_normalisation_and_redshift(dev):
    e = np.logspace(-1, 3, 4001)
    em = np.sqrt(e[1:] * e[:-1])
    k = int(np.argmin(abs(em - 1.0)))
    for shift in (1.0, 0.5, 0.25):
        y = rx.nthcomp(e, 60.0, 2.0, shift)[0]
        val = y[k] / (e[k + 1] - e[k]) * em[k] ** 2
        assert abs(val - 1.0) < 2e-3, (shift, val)
    z = 0.37
    a = rx.nthcomp(e, 30.0, 1.8, 1.0 / (1 + z))[0]
    b = rx.nthcomp(e * (1 + z), 30.0, 1.8, 1.0)[0]
    ok = (a > 0) & (b > 0)
    r = a[ok] / b[ok]
    assert ok.sum() > 3000 and np.ptp(r) / r.mean() < 1e-10


def test_bad_arguments(dev):
    e = log_grid(100, 0.1, 100.0)
    with pytest.raises(rx.RxError):
        rx.nthcomp(e, -5.0, 2.0)
    with pytest.raises(rx.RxError):
        rx.nthcomp(e[::-1].copy(), 60.0, 2.0)
