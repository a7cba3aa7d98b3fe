"""GPU evaluation on tables read by the FITS loaders (src/reltable.c:235-448):
relline and lamp-post tables at the released dimensions (25 spins x 30 mu0 x
100 radii x 40 g*, 20 spins x 250 heights x 100 radii) are written in the
reference's FITS layout by the independent writer of tests/helpers.py, loaded
with rx_tables_load_fits / rx_lptable_load_fits, and relline, relconv,
relline_lp and relconv_lp are compared with the FP64 oracle run on the arrays
that were written (not on what the library read)."""
import numpy as np
import pytest

import oracle
import relxill_amd as rx
from tests.helpers import (RELLINE_CFG0, log_grid, reflection_spectrum, rel_err, tables, write_fits_lptable,
                           write_fits_reltable)

pytestmark = pytest.mark.gpu


def _cfitsio_missing():
    st = rx.lib.rx_tables_load_fits(b"/nonexistent/rel_table_v0.5a.fits")
    return st != 0 and "cfitsio not found" in rx.lib.rx_last_error().decode()


@pytest.fixture(scope="module")
def fits_tabs(tmp_path_factory):
    if _cfitsio_missing():
        pytest.skip("cfitsio cannot be loaded here")
    tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    d = tmp_path_factory.mktemp("fits")
    # arrays of the released dims (synthetic fields), copied before the
    # resident tables are restored
    rx.tables_synthetic(17, 25, 30, 100, 40)
    rx.lptable_synthetic(19, 20, 250, 100)
    rel, lp = rx.tables_view()
    rel = {k: (np.array(v, copy=True) if hasattr(v, "shape") else v) for k, v in rel.items()}
    lp = {k: (np.array(v, copy=True) if hasattr(v, "shape") else v) for k, v in lp.items()}
    # signed angles in the file: the loader takes their absolute values (:370-377)
    lp_file = dict(lp)
    lp_file["del_"] = -lp["del_"]
    rx.tables_synthetic(1, 25, 22, 100, 20)
    rx.lptable_synthetic(3, 20, 100, 100)
    write_fits_reltable(d / "rel_table_v0.5a.fits", rel)
    write_fits_lptable(d / "rel_lp_table_v0.5b.fits", lp_file)
    assert rx.lib.rx_tables_load_fits(str(d / "rel_table_v0.5a.fits").encode()) == 0, rx.lib.rx_last_error()
    assert rx.lib.rx_lptable_load_fits(str(d / "rel_lp_table_v0.5b.fits").encode()) == 0, rx.lib.rx_last_error()
    yield oracle.Tables(rel, lp)
    rx.tables_synthetic(1, 25, 22, 100, 20)
    rx.lptable_synthetic(3, 20, 100, 100)


def test_models_on_fits_loaded_tables(fits_tabs):
    T = fits_tabs
    e = log_grid(3000, 0.1, 10.0)
    e_ref, refl = reflection_spectrum(n_incl=1)
    e_c = log_grid(2000, 0.1, 100.0)
    fin = oracle.rebin(e_c, e_ref, refl[0])
    cases = [(rx.MODEL_RELLINE, RELLINE_CFG0, e, None),
             (rx.MODEL_RELLINE, [6.4, 2.0, 4.0, 8.0, -0.6, 70.0, 4.0, 300.0, 0.1, 1.0], e, None),
             (rx.MODEL_RELCONV, [3.0, 3.0, 15.0, -0.4, 45.0, -1.0, 400.0, 0.0], e_c, fin),
             (rx.MODEL_RELLINELP, [6.4, 8.0, 0.9, 30.0, -1.0, 400.0, 0.0, 0.0, 2.0, 0.0], e, None),
             (rx.MODEL_RELLINELP, [6.4, 42.0, -0.3, 60.0, 7.0, 500.0, 0.2, 2.0, 2.4, 0.0], e, None),
             (rx.MODEL_RELCONVLP, [5.0, 0.3, 0.5, 35.0, -1.0, 400.0, 0.0, 2.0, 0.0], e_c, fin)]
    for model, p, grid, flux_in in cases:
        ref, st = oracle.eval_model(model, T, p, grid, flux_in=flux_in)
        assert st == 0
        for prec, tol in ((rx.FP64, 1e-8), (rx.FP32, 1e-4)):
            y, st = rx.eval_model(model, p, grid, flux_in=flux_in, precision=prec)
            err = rel_err(y, ref)
            print(f"FITS tables model {model} prec {prec}: max rel err {err:.3e}", flush=True)
            assert st == 0 and err < tol


def test_lamp_post_reflection_fractions_on_fits_tables(fits_tabs):
    """calc_refl_frac on the FITS-loaded lamp-post table (angles read as
    absolute values) equals the oracle's on the written arrays."""
    T = fits_tabs
    e = log_grid(300, 0.1, 100.0)
    reqs, refs = [], []
    for h, beta, a in ((4.0, 0.0, 0.9), (20.0, 0.4, -0.5), (70.0, 0.1, 0.2)):
        par = [h, beta, a, 30.0, -1.0, 400.0, 0.0, 2.0, 0.0]
        p, st = oracle.rel_params(oracle.MOD_RELCONVLP, par)
        sp, st2 = oracle.system_parameters(T, p)
        assert st == 0 and st2 == 0
        refs.append(sp["refl"])
        reqs.append(dict(model=rx.MODEL_RELCONVLP, energy=e, params=par, flux_in=np.ones(300)))
    b = rx.Batch(reqs)
    b.run(rx.FP64)
    b.sync()
    for i, r in enumerate(refs):
        got = b.reflfrac(i)
        assert np.max(np.abs(got / r - 1)) < 1e-12, (got, r)
    b.close()
