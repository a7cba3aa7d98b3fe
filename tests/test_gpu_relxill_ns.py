"""relxillNS on the GPU (the blackbody primary of k_primary and the blackbody
model's xillver table, interpolated in kTbb, A_Fe, logXi and logN) against
the oracle: lmodel.dat defaults, a randomized sweep over the lmodel.dat
ranges, a batch with all three primaries, and the XSPEC entry point.
Spectra are compared with conv_err (relative to the bin, or to 1e-5 of the
maximum below that: the FP64 round-off of the convolutions)."""
import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import conv_err, draw_relxill_params, log_grid, rel_err, tables

pytestmark = pytest.mark.gpu

TOL64 = 1e-8
TOL32 = 1e-4


@pytest.fixture(scope="module")
def ns():
    _, _, T = tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    rx.xilltable_synthetic(5, 7, 4, 6, 5, 10, 2999)
    rx.xillcptable_synthetic(6, 7, 4, 6, 5, 10, 2999)
    rx.xillnstable_synthetic(7, 8, 4, 6, 5, 10, 2999)
    yield (T, oracle.Xill(rx.xilltable_view()), oracle.Xill(rx.xilltable_view(nthcomp=True)),
           oracle.Xill(rx.xilltable_view(blackbody=True)))
    rx.xilltable_synthetic()
    rx.xillcptable_synthetic()
    rx.xillnstable_synthetic()


def draw_ns(rng):
    """relxillNS: Index1 Index2 Rbr a Incl Rin Rout z kTbb logxi Afe logN refl_frac"""
    return [rng.uniform(-1, 6), rng.uniform(-1, 6), rng.uniform(2, 100), rng.uniform(-0.998, 0.998),
            rng.uniform(5, 85), rng.choice([-1.0, rng.uniform(1.0, 30.0)]), rng.uniform(50, 1000),
            rng.choice([0.0, rng.uniform(0, 2)]), rng.uniform(0.5, 10.0), rng.uniform(1.0, 4.7),
            rng.uniform(0.5, 10.0), rng.uniform(15.0, 19.0), rng.choice([-1.0, rng.uniform(-5, 10)])]


def test_relxillns_default_parameters(ns):
    """lmodel.dat defaults on 3000 bins over 0.1-100 keV, FP64 and FP32, and
    the angle distribution."""
    T, _, _, xn = ns
    p = list(rx.DEFAULTS[rx.MODEL_RELXILLNS])
    e = log_grid(3000, 0.1, 100.0)
    ref, dist_ref, st = oracle.relxill_model(oracle.MOD_RELXILLNS, T, xn, p, e)
    assert st == 0
    for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
        b = rx.Batch([dict(model=rx.MODEL_RELXILLNS, energy=e, params=p)])
        b.run(prec)
        b.sync()
        y, st = b.output(0)
        d = b.dist(0, dist_ref.size)
        b.close()
        print(f"relxillNS defaults prec {prec}: rel {conv_err(y, ref):.3e} dist "
              f"{rel_err(d, dist_ref.reshape(-1)):.3e}", flush=True)
        assert st == 0
        assert conv_err(y, ref) < tol
        assert rel_err(d, dist_ref.reshape(-1)) < tol


def test_relxillns_random_sweep(ns):
    """10 parameter sets over the lmodel.dat ranges (kTbb 0.5-10 keV, logN
    15-19, both signs of refl_frac, redshift), one batch, FP64 and FP32."""
    T, _, _, xn = ns
    rng = np.random.default_rng(41)
    e = log_grid(1500, 0.1, 100.0)
    ps = [draw_ns(rng) for _ in range(10)]
    refs = [oracle.relxill_model(oracle.MOD_RELXILLNS, T, xn, p, e) for p in ps]
    assert sum(r[2] == 0 for r in refs) >= 6
    b = rx.Batch([dict(model=rx.MODEL_RELXILLNS, energy=e, params=p) for p in ps])
    try:
        for prec, tol in ((rx.FP64, TOL64), (rx.FP32, TOL32)):
            b.run(prec)
            b.sync()
            worst = 0.0
            for i, (ref, _, st) in enumerate(refs):
                y, st_g = b.output(i)
                assert (st_g == 0) == (st == 0), (i, ps[i], st, st_g)
                if st == 0:
                    worst = max(worst, conv_err(y, ref))
            print(f"relxillNS sweep prec {prec}: worst rel {worst:.3e}", flush=True)
            assert worst < tol
    finally:
        b.close()


def test_three_primaries_in_one_batch(ns):
    """relxill, relxillCp and relxillNS requests in one batch, each against
    the oracle with its own table (FP64)."""
    T, xt, xc, xn = ns
    rng = np.random.default_rng(42)
    e = log_grid(900, 0.1, 100.0)
    reqs, refs = [], []
    for _ in range(2):
        p = draw_relxill_params(rng)
        reqs.append(dict(model=rx.MODEL_RELXILL, energy=e, params=p))
        refs.append(oracle.relxill_model(oracle.MOD_RELXILL, T, xt, p, e))
        p = list(rx.DEFAULTS[rx.MODEL_RELXILLCP])
        p[12] = float(np.exp(rng.uniform(np.log(2.0), np.log(300.0))))
        reqs.append(dict(model=rx.MODEL_RELXILLCP, energy=e, params=p))
        refs.append(oracle.relxill_model(oracle.MOD_RELXILLCP, T, xc, p, e))
        p = draw_ns(rng)
        p[12] = 1.5
        reqs.append(dict(model=rx.MODEL_RELXILLNS, energy=e, params=p))
        refs.append(oracle.relxill_model(oracle.MOD_RELXILLNS, T, xn, p, e))
    b = rx.Batch(reqs)
    try:
        b.run(rx.FP64)
        b.sync()
        for i, ref in enumerate(refs):
            y, st = b.output(i)
            assert (st == 0) == (ref[2] == 0), (i, reqs[i]["model"], st, ref[2])
            if st == 0:
                err = conv_err(y, ref[0])
                print(f"request {i} model {reqs[i]['model']}: rel {err:.3e}", flush=True)
                assert err < TOL64
    finally:
        b.close()


def test_lmodrelxillns_entry_point(ns):
    """lmodrelxillns gives the FP64 result; rejected parameters give zeros
    and a status."""
    T, _, _, xn = ns
    e = log_grid(700, 0.2, 80.0)
    p = [3.0, 2.5, 20.0, 0.7, 40.0, -1.0, 400.0, 0.05, 2.4, 2.5, 3.0, 17.0, 2.0]
    ref, _, st = oracle.relxill_model(oracle.MOD_RELXILLNS, T, xn, p, e)
    assert st == 0
    y = rx.lmodrelxillns(e, p)
    print(f"lmodrelxillns: rel {conv_err(y, ref):.3e}", flush=True)
    assert rx.lib.rx_last_status() == 0 and conv_err(y, ref) < TOL64
    bad = list(p)
    bad[5], bad[6] = 50.0, 20.0  # Rin > Rout
    y = rx.lmodrelxillns(e, bad)
    assert not y.any() and rx.lib.rx_last_status() == rx.RX_ERR_PARAM
