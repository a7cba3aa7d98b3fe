"""The non-relativistic xillver models on the GPU (xillver, xillverCp,
xillverNS: k_xill with the inclination bracket as a two-bin angle
distribution, k_primary rebinning the table spectrum and adding the primary
on the caller grid; nthcomp from k_nth) against the oracle: lmodel.dat
defaults, randomized sweeps over the lmodel.dat ranges (parameters beyond
the tables too), a batch with the relativistic models, and the XSPEC entry
points."""
import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import draw_relxill_params, log_grid, rel_err, tables

pytestmark = pytest.mark.gpu

TOL = 1e-8  # no integration or convolution: FP64 in both precisions

MODELS = [(rx.MODEL_XILLVER, oracle.MOD_XILLVER, 0), (rx.MODEL_XILLVERCP, oracle.MOD_XILLVERCP, 1),
          (rx.MODEL_XILLVERNS, oracle.MOD_XILLVERNS, 2)]


@pytest.fixture(scope="module")
def xt():
    tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    rx.xilltable_synthetic(5, 7, 4, 6, 5, 10, 2999)
    rx.xillcptable_synthetic(6, 7, 4, 6, 5, 10, 2999)
    rx.xillnstable_synthetic(7, 8, 4, 6, 5, 10, 2999)
    yield [oracle.Xill(rx.xilltable_view()), oracle.Xill(rx.xilltable_view(nthcomp=True)),
           oracle.Xill(rx.xilltable_view(blackbody=True))]
    rx.xilltable_synthetic()
    rx.xillcptable_synthetic()
    rx.xillnstable_synthetic()


def draw(model, rng):
    rf = rng.choice([-1.0, rng.uniform(-5, 10)])
    z = rng.choice([0.0, rng.uniform(0, 2)])
    incl = rng.uniform(3, 89)
    if model == rx.MODEL_XILLVER:  # gamma Afe Ecut logxi z Incl refl_frac
        return [rng.uniform(0.9, 3.5), rng.uniform(0.3, 12.0), rng.uniform(3.0, 1200.0), rng.uniform(-0.5, 5.0), z,
                incl, rf]
    if model == rx.MODEL_XILLVERCP:  # gamma Afe kTe logxi logN z Incl refl_frac
        return [rng.uniform(1.2, 3.4), rng.uniform(0.5, 10.0), float(np.exp(rng.uniform(0.0, np.log(400.0)))),
                rng.uniform(0.0, 4.7), rng.choice([15.0, rng.uniform(15.0, 20.0)]), z, incl, rf]
    # kTbb Afe logN logxi z Incl refl_frac
    return [rng.uniform(0.4, 11.0), rng.uniform(0.5, 10.0), rng.uniform(15.0, 19.0), rng.uniform(1.0, 4.7), z,
            incl, rf]


@pytest.mark.parametrize("model,omodel,k", MODELS)
def test_default_parameters(xt, model, omodel, k):
    """lmodel.dat defaults on 3000 bins over 0.1-100 keV, both precisions,
    and with refl_frac 1 (the primary added)."""
    e = log_grid(3000, 0.1, 100.0)
    for rf in (None, 1.0):
        p = list(rx.DEFAULTS[model])
        if rf is not None:
            p[-1] = rf
        ref, st = oracle.xillver_model(omodel, xt[k], p, e)
        assert st == 0
        for prec in (rx.FP64, rx.FP32):
            b = rx.Batch([dict(model=model, energy=e, params=p)])
            b.run(prec)
            b.sync()
            y, st = b.output(0)
            b.close()
            print(f"model {model} refl_frac {p[-1]} prec {prec}: rel {rel_err(y, ref):.3e}", flush=True)
            assert st == 0 and rel_err(y, ref) < TOL


def test_random_sweeps_in_one_batch(xt):
    """8 sets per model over (and beyond) the lmodel.dat ranges, redshift,
    both signs of refl_frac, the three models and relxill in one batch."""
    _, _, T = tables()
    rng = np.random.default_rng(51)
    e = log_grid(1200, 0.1, 100.0)
    reqs, refs = [], []
    for model, omodel, k in MODELS:
        for _ in range(8):
            p = draw(model, rng)
            reqs.append(dict(model=model, energy=e, params=p))
            refs.append(oracle.xillver_model(omodel, xt[k], p, e))
    p = draw_relxill_params(rng)
    reqs.append(dict(model=rx.MODEL_RELXILL, energy=e, params=p))
    refs.append(oracle.relxill_model(oracle.MOD_RELXILL, T, xt[0], p, e)[::2])
    b = rx.Batch(reqs)
    try:
        b.run(rx.FP64)
        b.sync()
        worst = 0.0
        for i, (ref, st) in enumerate(refs):
            y, st_g = b.output(i)
            assert (st_g == 0) == (st == 0), (i, reqs[i]["params"], st, st_g)
            if st == 0:
                worst = max(worst, rel_err(y, ref))
        print(f"xillver models sweep: worst rel {worst:.3e}", flush=True)
        assert worst < TOL
    finally:
        b.close()


def test_entry_points(xt):
    """lmodxillver, lmodxillverdensnthcomp and lmodxillverns give the FP64
    result."""
    e = log_grid(700, 0.2, 80.0)
    for (model, omodel, k), fn, p in zip(
            MODELS, (rx.lmodxillver, rx.lmodxillverdensnthcomp, rx.lmodxillverns),
            ([2.3, 2.0, 150.0, 2.5, 0.05, 45.0, 1.5], [2.1, 3.0, 35.0, 2.0, 16.5, 0.05, 45.0, 1.5],
             [1.6, 2.0, 17.0, 2.5, 0.05, 45.0, 1.5])):
        ref, st = oracle.xillver_model(omodel, xt[k], p, e)
        assert st == 0
        y = fn(e, p)
        print(f"{fn.__name__}: rel {rel_err(y, ref):.3e}", flush=True)
        assert rx.lib.rx_last_status() == 0 and rel_err(y, ref) < TOL
