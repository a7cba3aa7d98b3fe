"""C-ABI checks that need no GPU: the library loads, exports every function
include/relxill_gpu.h declares, its host-side parameter processing equals the
oracle's get_rel_params restatement (src/ModelDefinition.cpp:168-334), and the
synthetic tables satisfy the grid invariants the interpolation assumes
(src/Relprofile.cpp:181-184, 195-238)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
import relxill_amd as rx
import relxill_amd.capi as capi
from tests.helpers import draw_wide_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "relxill_gpu.h")


def declared_functions():
    """Names of the functions declared in the header (comments stripped)."""
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"^\s*#.*$", "", src, flags=re.M)
    src = re.sub(r"typedef\s+struct\s*\{.*?\}\s*\w+\s*;", "", src, flags=re.S)
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src)
    return sorted(set(n for n in names if n not in ("if", "while", "return", "sizeof")))


def test_header_parses_to_the_known_entry_points():
    names = declared_functions()
    for must in ("lmodrelline", "lmodrellinelp", "lmodrelconv", "lmodrelconvlp", "rx_eval_model",
                 "rx_batch_create", "rx_batch_run", "rx_server_start", "rx_server_submit"):
        assert must in names
    assert len(names) >= 40, names


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_only_declared_functions():
    declared = set(declared_functions())
    bound = set(capi._EXPORTS)
    assert bound <= declared, bound - declared
    # the Python mirror covers the whole ABI
    assert declared <= bound, declared - bound


@pytest.mark.parametrize("model", [1, 2, 11, 12])
def test_model_params_equal_oracle(model):
    rng = np.random.default_rng(100 + model)
    n_ok = n_rej = 0
    for _ in range(400):
        par = draw_wide_params(model, rng)
        out, st = rx.model_params(model, par)
        rp, st_o = oracle.rel_params(model, par)
        assert st == st_o, (par, st, st_o)
        if st != 0:
            n_rej += 1
            continue
        n_ok += 1
        ref = [rp.a, rp.incl, rp.emis1, rp.emis2, rp.rbr, rp.rin, rp.rout, rp.lineE, rp.z, rp.height,
               rp.gamma, rp.beta]
        if rp.emis_type == 2:  # lamp post: the broken power law fields are unused
            out, ref = np.r_[out[:2], out[5:]], np.r_[ref[:2], ref[5:]]
        assert np.array_equal(out, np.asarray(ref)), (par, out, ref)
    assert n_ok > 50


def test_parameter_bounds_of_the_reference():
    """check_parameter_bounds (src/ModelDefinition.cpp:168-210): Rin < ISCO is
    reset to the ISCO, Rout <= Rin and |a| > 0.9982 or incl outside 3-87 deg
    are rejected; returning radiation is accepted for relline_lp, relconv_lp
    and relxilllp (switch 1, -1, 2), other switch values rejected."""
    out, st = rx.model_params(1, [6.4, 3, 3, 15, 0.998, 30, 1.0, 400, 0, 0])
    assert st == 0 and out[5] == oracle.lib.orc_kerr_rms(0.998)
    assert rx.model_params(11, [3, 3, 15, 0.5, 30, 50.0, 40.0, 0])[1] == 1
    assert rx.model_params(11, [3, 3, 15, 0.999, 30, -1, 400, 0])[1] == 1
    assert rx.model_params(11, [3, 3, 15, 0.5, 2.0, -1, 400, 0])[1] == 1
    assert rx.model_params(11, [3, 3, 15, 0.5, 88.0, -1, 400, 0])[1] == 1
    assert rx.model_params(2, [6.4, 6.0, 0.998, 30, -1, 400, 0, 0, 2, 1])[1] == 0
    assert rx.model_params(2, [6.4, 6.0, 0.998, 30, -1, 400, 0, 0, 2, 3])[1] == 1
    assert rx.model_params(-2, [6.0, 0.0, 0.998, 30, -1, 400, 0, 2, 3.1, 1, 300, 1, 1, 0])[1] == 0
    assert rx.model_params(-2, [6.0, 0.0, 0.998, 30, -1, 400, 0, 2, 3.1, 1, 300, 1, 3, 0])[1] == 1


def test_synthetic_table_invariants():
    rx.tables_synthetic(1, 25, 22, 100, 20)
    rel, lp = rx.tables_view()
    na, nm, nr, ng = rel["n_a"], rel["n_mu0"], rel["n_r"], rel["n_g"]
    a, mu0 = rel["a"], rel["mu0"]
    assert np.all(np.diff(a) > 0) and a[0] <= -0.998 and a[-1] >= 0.998
    assert np.all(np.diff(mu0) > 0) and mu0[0] <= np.cos(np.radians(87)) and mu0[-1] >= np.cos(np.radians(3))
    r = rel["r"].reshape(na, nm, nr)
    assert np.all(np.diff(r, axis=2) < 0)  # descending radii
    # outermost radius within 1% below 1000 rg: the 1000 rg snap applies
    assert np.all((r[..., 0] < 1000.0) & (r[..., 0] * 1.01 > 1000.0))
    for ia in range(na):
        assert abs(r[ia, 0, -1] - oracle.lib.orc_kerr_rms(float(a[ia]))) < 1e-3 * r[ia, 0, -1]
    gmin, gmax = rel["gmin"], rel["gmax"]
    assert np.all(gmin > 0) and np.all(gmax > gmin)
    t1, t2 = rel["trff1"], rel["trff2"]
    assert np.all(t1 >= 0) and np.all(t2 >= 0)
    c1, c2 = rel["cosne1"], rel["cosne2"]
    assert np.all((c1 >= 0) & (c1 <= 1)) and np.all((c2 >= 0) & (c2 <= 1))
    # lamp-post table: heights ascending per spin, radii ascending to 1000
    rx.lptable_synthetic(3, 20, 100, 100)
    rel, lp = rx.tables_view()
    h = lp["h"].reshape(lp["n_a"], lp["n_h"])
    assert np.all(np.diff(h, axis=1) > 0)
    rad = lp["rad"].reshape(lp["n_a"], lp["n_r"])
    assert np.all(np.diff(rad, axis=1) > 0) and np.allclose(rad[:, -1], 1000.0, rtol=1e-6)
    assert np.all(lp["intens"] > 0)


def test_table_file_round_trip(tmp_path):
    rx.tables_synthetic(7, 6, 5, 30, 12)
    rel0, lp0 = rx.tables_view()
    p = str(tmp_path / "t.rxtb").encode()
    assert rx.lib.rx_tables_save(p) == 0
    rx.tables_synthetic(1, 25, 22, 100, 20)
    assert rx.lib.rx_tables_load(p) == 0
    rel1, _ = rx.tables_view()
    for k in ("a", "mu0", "r", "gmin", "gmax", "trff1", "trff2", "cosne1", "cosne2"):
        assert np.array_equal(rel0[k], rel1[k]), k
    rx.tables_synthetic(1, 25, 22, 100, 20)  # restore the default tables


def _fresh_process(code, table_path):
    import subprocess
    import sys
    env = dict(os.environ, RELXILL_TABLE_PATH=str(table_path))
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    return out.stdout


def test_table_path_without_tables_is_an_error(tmp_path):
    """RELXILL_TABLE_PATH set but holding no tables: the reference cannot open
    its table and fails (src/reltable.c:255-268), so the library returns
    RX_ERR_TABLE with the reason and zero flux instead of evaluating on
    synthetic tables; explicitly loaded tables clear the error."""
    code = ("import numpy as np, relxill_amd as rx, relxill_amd.capi\n"
            "e = np.geomspace(0.1, 10.0, 101)\n"
            "f, st = rx.eval_model(rx.MODEL_RELLINE, [6.4, 3, 3, 15, 0.9, 30, -1, 400, 0, 0], e,"
            " raise_on_error=False)\n"
            "print(st, int(np.any(f != 0)))\n"
            "print(rx.lib.rx_last_error().decode())\n"
            "rel, lp = rx.capi._RelView(), rx.capi._LpView()\n"
            "import ctypes\n"
            "print(rx.lib.rx_tables_view(ctypes.byref(rel), None))\n"
            "rx.tables_synthetic(1, 5, 4, 20, 8)\n"
            "print(rx.lib.rx_tables_view(ctypes.byref(rel), None), rel.n_a)\n"
            "print(rx.lib.rx_tables_view(None, ctypes.byref(lp)))\n"
            "print(rx.lib.rx_last_error().decode())\n")
    lines = _fresh_process(code, tmp_path).splitlines()
    assert lines[0] == "10 0"
    assert "rel_table_v0.5a.fits" in lines[1] and str(tmp_path) in lines[1]
    assert lines[2] == "10"
    assert lines[3] == "0 5"
    assert lines[4] == "10" and "rel_lp_table_v0.5b.fits" in lines[5]


def test_unreadable_table_file_is_an_error(tmp_path):
    """A table file that exists but cannot be read is reported with the
    loader's message (never replaced by synthetic data)."""
    (tmp_path / "rel_table.rxtb").write_bytes(b"RXTB but truncated")
    code = ("import ctypes, relxill_amd as rx, relxill_amd.capi\n"
            "rel = rx.capi._RelView()\n"
            "print(rx.lib.rx_tables_view(ctypes.byref(rel), None))\n"
            "print(rx.lib.rx_last_error().decode())\n")
    lines = _fresh_process(code, tmp_path).splitlines()
    assert lines[0] == "10"
    assert "rel_table.rxtb" in lines[1]
