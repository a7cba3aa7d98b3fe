"""Loaders of the reference's FITS tables (src/reltable.c:121-448): files laid
out as rel_table_v0.5a.fits / rel_lp_table_v0.5b.fits, written here by an
independent writer (tests/helpers.py), load into exactly the arrays they were
written from; the lamp-post angles come back as absolute values (:370-377);
malformed files are rejected with RX_ERR_TABLE.  No GPU needed."""
import os
import subprocess
import sys

import numpy as np
import pytest

import relxill_amd as rx
from tests.helpers import write_fits_lptable, write_fits_reltable

REL_KEYS = ("a", "mu0", "r", "gmin", "gmax", "trff1", "trff2", "cosne1", "cosne2")
LP_KEYS = ("a", "h", "rad", "intens", "del_", "del_inc")


def _cfitsio_missing():
    st = rx.lib.rx_tables_load_fits(b"/nonexistent/rel_table_v0.5a.fits")
    return st != 0 and "cfitsio not found" in rx.lib.rx_last_error().decode()


pytestmark = pytest.mark.skipif(_cfitsio_missing(), reason="cfitsio cannot be loaded here")


def _copy(view):
    return {k: (np.array(v, copy=True) if hasattr(v, "shape") else v) for k, v in view.items()}


@pytest.fixture
def restore_tables():
    yield
    rx.tables_synthetic(1, 25, 22, 100, 20)
    rx.lptable_synthetic(3, 20, 100, 100)


def test_relline_fits_round_trip(tmp_path, restore_tables):
    rx.tables_synthetic(7, 6, 5, 30, 12)
    rel0, _ = rx.tables_view()
    rel0 = _copy(rel0)
    path = tmp_path / "rel_table_v0.5a.fits"
    write_fits_reltable(path, rel0)
    rx.tables_synthetic(1, 25, 22, 100, 20)
    assert rx.lib.rx_tables_load_fits(str(path).encode()) == 0, rx.lib.rx_last_error()
    rel1, _ = rx.tables_view()
    assert (rel1["n_a"], rel1["n_mu0"], rel1["n_r"], rel1["n_g"]) == (6, 5, 30, 12)
    for k in REL_KEYS:
        assert np.array_equal(rel0[k], rel1[k]), k


def test_relline_fits_at_the_reference_dims(tmp_path, restore_tables):
    """The released table's dims (src/common.h:69-74): 25 spins, 30 mu0, 100
    radii, 40 g* points."""
    rx.tables_synthetic(11, 25, 30, 100, 40)
    rel0, _ = rx.tables_view()
    rel0 = _copy(rel0)
    path = tmp_path / "rel_table_v0.5a.fits"
    write_fits_reltable(path, rel0)
    rx.tables_synthetic(1, 25, 22, 100, 20)
    assert rx.lib.rx_tables_load_fits(str(path).encode()) == 0, rx.lib.rx_last_error()
    rel1, _ = rx.tables_view()
    for k in REL_KEYS:
        assert np.array_equal(rel0[k], rel1[k]), k


def test_lamp_post_fits_round_trip_with_signed_angles(tmp_path, restore_tables):
    rx.lptable_synthetic(5, 4, 7, 9)
    _, lp0 = rx.tables_view()
    lp0 = _copy(lp0)
    signed = dict(lp0)
    rng = np.random.default_rng(3)
    flip = rng.random(lp0["del_"].shape) < 0.5
    signed["del_"] = np.where(flip, -lp0["del_"], lp0["del_"]).astype(np.float32)
    signed["del_inc"] = np.where(~flip, -lp0["del_inc"], lp0["del_inc"]).astype(np.float32)
    path = tmp_path / "rel_lp_table_v0.5b.fits"
    write_fits_lptable(path, signed)
    rx.lptable_synthetic(3, 20, 100, 100)
    assert rx.lib.rx_lptable_load_fits(str(path).encode()) == 0, rx.lib.rx_last_error()
    _, lp1 = rx.tables_view()
    assert (lp1["n_a"], lp1["n_h"], lp1["n_r"]) == (4, 7, 9)
    for k in LP_KEYS:
        assert np.array_equal(lp0[k], lp1[k]), k  # |del| and |del_inc| restored


def test_malformed_fits_files_are_rejected(tmp_path, restore_tables):
    from tests import helpers
    rx.tables_synthetic(7, 3, 3, 10, 6)
    rel, _ = rx.tables_view()
    rel = _copy(rel)
    node = lambda i, j, nr=10: helpers._fits_bintable(f"{i + 1}_{j + 1}", [
        (k, rel[k].reshape(3, 3, 10)[i, j, :nr]) for k in ("r", "gmin", "gmax")] + [
        (k, rel[k].reshape(3, 3, 10, 6)[i, j, :nr]) for k in ("trff1", "trff2", "cosne1", "cosne2")])
    head = [helpers._fits_primary(), helpers._fits_bintable("a", [("a", rel["a"])]),
            helpers._fits_bintable("mu0", [("mu0", rel["mu0"])])]

    def write(name, blocks):
        with open(tmp_path / name, "wb") as f:
            for blk in blocks:
                f.write(blk)
        return str(tmp_path / name).encode()

    # missing file
    assert rx.lib.rx_tables_load_fits(str(tmp_path / "none.fits").encode()) == rx.RX_ERR_TABLE
    assert "cannot open" in rx.lib.rx_last_error().decode()
    # the spin axis has 3 entries but the file ends after the nodes of 2 spins
    p = write("short.fits", head + [node(i, j) for i in range(2) for j in range(3)])
    assert rx.lib.rx_tables_load_fits(p) == rx.RX_ERR_TABLE
    assert "HDU 10" in rx.lib.rx_last_error().decode()
    # one node with fewer radii than the first
    p = write("ragged.fits", head + [node(i, j, 9 if (i, j) == (1, 2) else 10) for i in range(3) for j in range(3)])
    assert rx.lib.rx_tables_load_fits(p) == rx.RX_ERR_TABLE
    assert "2_3 has other dimensions" in rx.lib.rx_last_error().decode()
    # a spin axis that is not ascending fails the grid checks
    bad = dict(rel)
    bad["a"] = rel["a"][::-1].copy()
    write_fits_reltable(tmp_path / "bad.fits", bad)
    assert rx.lib.rx_tables_load_fits(str(tmp_path / "bad.fits").encode()) == rx.RX_ERR_TABLE
    assert "ascending" in rx.lib.rx_last_error().decode()
    # failed loads leave the resident table untouched
    now, _ = rx.tables_view()
    assert now["n_a"] == 3 and np.array_equal(now["a"], rel["a"])


def test_tables_found_under_the_reference_file_names(tmp_path):
    """RELXILL_TABLE_PATH with rel_table_v0.5a.fits / rel_lp_table_v0.5b.fits
    (the reference's names, src/common.h:73,80): a fresh process loads them."""
    rx.tables_synthetic(9, 5, 4, 20, 8)
    rx.lptable_synthetic(9, 3, 5, 6)
    rel, lp = rx.tables_view()
    write_fits_reltable(tmp_path / "rel_table_v0.5a.fits", _copy(rel))
    write_fits_lptable(tmp_path / "rel_lp_table_v0.5b.fits", _copy(lp))
    rx.tables_synthetic(1, 25, 22, 100, 20)
    rx.lptable_synthetic(3, 20, 100, 100)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import relxill_amd as rx\n"
            "rel, lp = rx.tables_view()\n"
            "print(rel['n_a'], rel['n_mu0'], rel['n_r'], rel['n_g'], lp['n_a'], lp['n_h'], lp['n_r'])\n")
    env = dict(os.environ, RELXILL_TABLE_PATH=str(tmp_path))
    out = subprocess.run([sys.executable, "-c", code], cwd=root, env=env, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split() == ["5", "4", "20", "8", "3", "5", "6"]
