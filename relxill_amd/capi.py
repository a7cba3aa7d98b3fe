"""ctypes mirror of include/relxill_gpu.h.

Function names, argument meaning and error behaviour follow the XSPEC local
models of the reference (src/create_wrapper_xspec.py:153-162,
src/LocalModel.cpp:32-160): energy grids are n+1 ascending edges in keV,
fluxes n values in photons/bin; convolution models take the input spectrum in
``flux`` and overwrite it.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "librelxill_gpu.so")

RX_OK = 0
RX_ERR_PARAM = 1
RX_ERR_UNSUPPORTED = 2
RX_ERR_TABLE = 10
RX_ERR_INPUT = 50
RX_ERR_DEVICE = 100
RX_ERR_ARG = 101

MODEL_RELLINE = 1
MODEL_RELLINELP = 2
MODEL_RELCONV = 11
MODEL_RELCONVLP = 12
MODEL_RELXILL = -1
MODEL_RELXILLLP = -2
MODEL_RELXILL_CONV = -1001  # project id: relxill-style conv of a resident reflection set
MODEL_RELXILLCP = -1010  # project id: relxillCp (relxill with the nthcomp primary)
MODEL_RELXILLLPCP = -1011  # project id: relxilllpCp
MODEL_RELXILLNS = -1012  # project id: relxillNS (relxill with the blackbody primary)
MODEL_XILLVER = 0
MODEL_XILLVERCP = 100
MODEL_XILLVERNS = -101

FP64 = 0
FP32 = 1

N_PARAMS = {MODEL_RELLINE: 10, MODEL_RELCONV: 8, MODEL_RELLINELP: 10, MODEL_RELCONVLP: 9,
            MODEL_RELXILL_CONV: 8, MODEL_RELXILL: 13, MODEL_RELXILLLP: 14, MODEL_RELXILLCP: 14,
            MODEL_RELXILLLPCP: 17, MODEL_RELXILLNS: 13, MODEL_XILLVER: 7, MODEL_XILLVERCP: 8,
            MODEL_XILLVERNS: 7}

# default parameter values of the reference's lmodel.dat
# (src/modelfiles/lmodel_relxill_public.dat:1-59)
DEFAULTS = {
    MODEL_RELXILL: [3.0, 3.0, 15.0, 0.998, 30.0, -1.0, 400.0, 0.0, 2.0, 3.1, 1.0, 300.0, 3.0],
    MODEL_RELXILLLP: [6.0, 0.0, 0.998, 30.0, -1.0, 400.0, 0.0, 2.0, 3.1, 1.0, 300.0, 1.0, 1.0, 0.0],
    MODEL_RELLINE: [6.4, 3.0, 3.0, 15.0, 0.998, 30.0, -1.0, 400.0, 0.0, 0.0],
    MODEL_RELCONV: [3.0, 3.0, 15.0, 0.998, 30.0, -1.0, 400.0, 0.0],
    MODEL_RELLINELP: [6.4, 6.0, 0.998, 30.0, -1.0, 400.0, 0.0, 0.0, 2.0, 1.0],
    MODEL_RELCONVLP: [6.0, 0.0, 0.998, 30.0, -1.0, 400.0, 0.0, 2.0, 1.0],
    MODEL_RELXILL_CONV: [3.0, 3.0, 15.0, 0.998, 30.0, -1.0, 400.0, 0.0],
    # lmodel_relxill_public.dat:96-129
    MODEL_RELXILLCP: [30.0, 0.998, -1.0, 400.0, 15.0, 3.0, 3.0, 0.0, 2.0, 3.1, 15.0, 1.0, 60.0, 3.0],
    MODEL_RELXILLLPCP: [30.0, 0.998, -1.0, 400.0, 6.0, 0.0, 2.0, 3.1, 15.0, 1.0, 60.0, 1.0, 0.0, 0.0, 0.0, 1.0,
                        0.0],
    # lmodel_relxill_public.dat:140-153
    MODEL_RELXILLNS: [3.0, 3.0, 15.0, 0.998, 30.0, -1.0, 400.0, 0.0, 2.0, 3.1, 1.0, 15.0, 3.0],
    # lmodel_relxill_public.dat:77-138
    MODEL_XILLVER: [2.0, 1.0, 300.0, 3.1, 0.0, 30.0, -1.0],
    MODEL_XILLVERCP: [2.0, 1.0, 60.0, 3.1, 15.0, 0.0, 30.0, -1.0],
    MODEL_XILLVERNS: [2.0, 1.0, 15.0, 3.1, 0.0, 30.0, -1.0],
}


class RxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"relxill_gpu status {status}: {msg}")
        self.status = status


_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)


class _Request(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int), ("n_flux", ctypes.c_int), ("energy", _D),
                ("parameter", _D), ("flux_in", _D), ("refl_set", ctypes.c_int)]


class _RelView(ctypes.Structure):
    _fields_ = [("n_a", ctypes.c_int), ("n_mu0", ctypes.c_int), ("n_r", ctypes.c_int),
                ("n_g", ctypes.c_int), ("a", _F), ("mu0", _F), ("r", _F), ("gmin", _F),
                ("gmax", _F), ("trff1", _F), ("trff2", _F), ("cosne1", _F), ("cosne2", _F)]


class _XillView(ctypes.Structure):
    _fields_ = [("n_par", ctypes.c_int), ("pid", ctypes.c_int * 6), ("n_val", ctypes.c_int * 6),
                ("val", _F * 6), ("n_ener", ctypes.c_int), ("elo", _F), ("ehi", _F), ("spec", _F)]


class _RradView(ctypes.Structure):
    _fields_ = [("n_spin", ctypes.c_int), ("n_rad", ctypes.c_int), ("n_g", ctypes.c_int)] + \
               [(k, _D) for k in ("spin", "rlo", "rhi", "tf_r", "gmin", "gmax", "frac_g")]


class _LpView(ctypes.Structure):
    _fields_ = [("n_a", ctypes.c_int), ("n_h", ctypes.c_int), ("n_r", ctypes.c_int),
                ("a", _F), ("h", _F), ("rad", _F), ("intens", _F), ("del_", _F),
                ("del_inc", _F)]


_EXPORTS = {
    # name: (restype, argtypes)
    "lmodrelline": (None, [_D, ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_char_p]),
    "lmodrellinelp": (None, [_D, ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_char_p]),
    "lmodrelconv": (None, [_D, ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_char_p]),
    "lmodrelconvlp": (None, [_D, ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_char_p]),
    "lmodrelxill": (None, [_D, ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_char_p]),
    "lmodrelxilllp": (None, [_D, ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_char_p]),
    "rx_xilltable_synthetic": (ctypes.c_int, [ctypes.c_uint64] + [ctypes.c_int] * 6),
    "rx_xilltable_load_fits": (ctypes.c_int, [ctypes.c_char_p]),
    "rx_xilltable_get_view": (ctypes.c_int, [ctypes.POINTER(_XillView)]),
    "rx_xillcptable_synthetic": (ctypes.c_int, [ctypes.c_uint64] + [ctypes.c_int] * 6),
    "rx_xillcptable_synthetic_dens": (ctypes.c_int, [ctypes.c_uint64] + [ctypes.c_int] * 7),
    "rx_xillcptable_load_fits": (ctypes.c_int, [ctypes.c_char_p]),
    "rx_xillcptable_get_view": (ctypes.c_int, [ctypes.POINTER(_XillView)]),
    "lmodrelxilldensnthcomp": (None, [_D, ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_char_p]),
    "lmodrelxilllpdensnthcomp": (None, [_D, ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_char_p]),
    "lmodrelxillns": (None, [_D, ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_char_p]),
    "lmodxillver": (None, [_D, ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_char_p]),
    "lmodxillverdensnthcomp": (None, [_D, ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_char_p]),
    "lmodxillverns": (None, [_D, ctypes.c_int, _D, ctypes.c_int, _D, _D, ctypes.c_char_p]),
    "rx_xillnstable_synthetic": (ctypes.c_int, [ctypes.c_uint64] + [ctypes.c_int] * 6),
    "rx_xillnstable_load_fits": (ctypes.c_int, [ctypes.c_char_p]),
    "rx_xillnstable_get_view": (ctypes.c_int, [ctypes.POINTER(_XillView)]),
    "rx_rradtable_synthetic": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "rx_rradtable_load_fits": (ctypes.c_int, [ctypes.c_char_p]),
    "rx_rradtable_get_view": (ctypes.c_int, [ctypes.POINTER(_RradView)]),
    "rx_eval_model": (ctypes.c_int, [ctypes.c_int, _D, _D, _D, ctypes.c_int, ctypes.c_int]),
    "rx_last_status": (ctypes.c_int, []),
    "rx_last_error": (ctypes.c_char_p, []),
    "rx_model_params": (ctypes.c_int, [ctypes.c_int, _D, _D]),
    "rx_init": (ctypes.c_int, [ctypes.c_int]),
    "rx_shutdown": (None, []),
    "rx_tables_synthetic": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int]),
    "rx_lptable_synthetic": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int]),
    "rx_tables_load": (ctypes.c_int, [ctypes.c_char_p]),
    "rx_lptable_load": (ctypes.c_int, [ctypes.c_char_p]),
    "rx_tables_save": (ctypes.c_int, [ctypes.c_char_p]),
    "rx_tables_load_fits": (ctypes.c_int, [ctypes.c_char_p]),
    "rx_lptable_load_fits": (ctypes.c_int, [ctypes.c_char_p]),
    "rx_lptable_save": (ctypes.c_int, [ctypes.c_char_p]),
    "rx_tables_view": (ctypes.c_int, [ctypes.POINTER(_RelView), ctypes.POINTER(_LpView)]),
    "rx_set_conv_grid": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "rx_set_nr_fine": (ctypes.c_int, [ctypes.c_int]),
    "rx_set_fp32_guard": (ctypes.c_int, [ctypes.c_double]),
    "rx_batch_create": (ctypes.c_void_p, [ctypes.POINTER(_Request), ctypes.c_int]),
    "rx_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "rx_batch_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "rx_batch_output": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D]),
    "rx_batch_status": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "rx_batch_dist": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int]),
    "rx_batch_reflfrac": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D]),
    "rx_batch_zones": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int]),
    "rx_batch_set_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "rx_reflsets_clear": (ctypes.c_int, []),
    "rx_batch_destroy": (None, [ctypes.c_void_p]),
    "rx_reflset_create": (ctypes.c_int, [_D, ctypes.c_int, _D, ctypes.c_int]),
    "rx_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "rx_device_synchronize": (ctypes.c_int, []),
    "rx_copy_test": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, _D]),
    "rx_nthcomp": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D]),
    "rx_batch_kernel_ms": (ctypes.c_int, [ctypes.c_void_p, _D]),
    "rx_batch_counters": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "rx_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "rx_eval_model_noalloc": (ctypes.c_int, [ctypes.c_int, _D, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "rx_set_graphs": (ctypes.c_int, [ctypes.c_int]),
    "rx_batch_graph_replays": (ctypes.c_int64, [ctypes.c_void_p]),
    "rx_server_start": (ctypes.c_void_p, [ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int]),
    "rx_server_submit": (ctypes.c_int64, [ctypes.c_void_p, ctypes.POINTER(_Request)]),
    "rx_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, _D,
                                      ctypes.POINTER(ctypes.c_int), _D]),
    "rx_server_replay": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_Request), ctypes.c_int, _D,
                                        ctypes.POINTER(ctypes.c_int64)]),
    "rx_server_drain": (ctypes.c_int, [ctypes.c_void_p]),
    "rx_server_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "rx_server_stop": (None, [ctypes.c_void_p]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m relxill_amd.build` "
            "(there is no CPU fallback)")
    lib_ = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _EXPORTS.items():
        fn = getattr(lib_, name)
        fn.restype = res
        fn.argtypes = args
    return lib_


lib = _load()


def _dp(a):
    return a.ctypes.data_as(_D)


def _arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _check(status):
    if status != RX_OK:
        raise RxError(status, lib.rx_last_error().decode())


def model_params(model, params):
    """Processed parameters: a, incl[rad], Index1, Index2, Rbr, Rin, Rout, lineE, z, h, gamma, beta."""
    p = _arr(params)
    out = np.zeros(12)
    st = lib.rx_model_params(model, _dp(p), _dp(out))
    return out, st


def eval_model(model, params, energy, flux_in=None, precision=FP64, raise_on_error=True):
    """One evaluation through rx_eval_model; returns (flux, status)."""
    e = _arr(energy)
    p = _arr(params)
    n = e.size - 1
    flux = _arr(flux_in).copy() if flux_in is not None else np.zeros(n)
    st = lib.rx_eval_model(model, _dp(p), _dp(e), _dp(flux), n, precision)
    if raise_on_error and st not in (RX_OK,):
        raise RxError(st, lib.rx_last_error().decode())
    return flux, st


def _lmod(fn, energy, params, flux=None):
    e = _arr(energy)
    p = _arr(params)
    n = e.size - 1
    f = _arr(flux).copy() if flux is not None else np.zeros(n)
    fn(_dp(e), n, _dp(p), 0, _dp(f), None, None)
    return f


def lmodrelline(energy, params):
    return _lmod(lib.lmodrelline, energy, params)


def lmodrellinelp(energy, params):
    return _lmod(lib.lmodrellinelp, energy, params)


def nthcomp(energy, kte, gam, ener_shift=1.0):
    """rx_nthcomp: nthcomp spectra (photons per bin) on one grid for arrays
    (or scalars) of kTe, Gamma and energy shift; returns (n_spec, n_bins)."""
    e = np.ascontiguousarray(energy, dtype=np.float64)
    kte = np.atleast_1d(np.asarray(kte, dtype=np.float64))
    n = kte.size
    gam = np.ascontiguousarray(np.broadcast_to(np.asarray(gam, dtype=np.float64), (n,)))
    sh = np.ascontiguousarray(np.broadcast_to(np.asarray(ener_shift, dtype=np.float64), (n,)))
    kte = np.ascontiguousarray(kte)
    out = np.zeros((n, e.size - 1))
    _check(lib.rx_nthcomp(e.ctypes.data_as(_D), e.size - 1, n, kte.ctypes.data_as(_D), gam.ctypes.data_as(_D),
                          sh.ctypes.data_as(_D), out.ctypes.data_as(_D)))
    return out


def lmodrelconv(energy, params, flux):
    return _lmod(lib.lmodrelconv, energy, params, flux)


def lmodrelconvlp(energy, params, flux):
    return _lmod(lib.lmodrelconvlp, energy, params, flux)


def lmodrelxill(energy, params):
    return _lmod(lib.lmodrelxill, energy, params)


def lmodrelxilllp(energy, params):
    return _lmod(lib.lmodrelxilllp, energy, params)


def lmodrelxilldensnthcomp(energy, params):
    return _lmod(lib.lmodrelxilldensnthcomp, energy, params)


def lmodrelxilllpdensnthcomp(energy, params):
    return _lmod(lib.lmodrelxilllpdensnthcomp, energy, params)


def lmodrelxillns(energy, params):
    return _lmod(lib.lmodrelxillns, energy, params)


def lmodxillver(energy, params):
    return _lmod(lib.lmodxillver, energy, params)


def lmodxillverdensnthcomp(energy, params):
    return _lmod(lib.lmodxillverdensnthcomp, energy, params)


def lmodxillverns(energy, params):
    return _lmod(lib.lmodxillverns, energy, params)


def xilltable_synthetic(seed=5, n_gam=13, n_afe=5, n_lxi=8, n_ect=6, n_incl=10, n_ener=2999):
    _check(lib.rx_xilltable_synthetic(seed, n_gam, n_afe, n_lxi, n_ect, n_incl, n_ener))


def xillcptable_synthetic(seed=6, n_gam=13, n_afe=5, n_lxi=8, n_kte=6, n_incl=10, n_ener=2999):
    """Synthetic table of the nthcomp models (fourth axis kTe, 1-400 keV)."""
    _check(lib.rx_xillcptable_synthetic(seed, n_gam, n_afe, n_lxi, n_kte, n_incl, n_ener))


def xillcptable_synthetic_dens(seed=8, n_dens=5, n_gam=13, n_afe=5, n_lxi=8, n_kte=6, n_incl=10, n_ener=2999):
    """Six-parameter synthetic table of the nthcomp models: Gamma, A_Fe,
    logXi, kTe, Dens (15-19), Incl (Gamma interpolated outside the others)."""
    _check(lib.rx_xillcptable_synthetic_dens(seed, n_dens, n_gam, n_afe, n_lxi, n_kte, n_incl, n_ener))


def xillnstable_synthetic(seed=7, n_ktbb=12, n_afe=5, n_lxi=8, n_dens=5, n_incl=10, n_ener=2999):
    """Synthetic table of the blackbody model (kTbb, A_Fe, logXi, Dens, Incl)."""
    _check(lib.rx_xillnstable_synthetic(seed, n_ktbb, n_afe, n_lxi, n_dens, n_incl, n_ener))


def xilltable_view(nthcomp=False, blackbody=False):
    """Copy of a resident xillver table (the nthcomp models' with
    nthcomp=True, the blackbody model's with blackbody=True): dict with n_par,
    pid, n_val, val (list of arrays), n_ener, elo, ehi, spec (rows x n_ener, as
    stored in the file)."""
    v = _XillView()
    get = (lib.rx_xillnstable_get_view if blackbody else
           lib.rx_xillcptable_get_view if nthcomp else lib.rx_xilltable_get_view)
    _check(get(ctypes.byref(v)))
    f = lambda ptr, n: np.ctypeslib.as_array(ptr, shape=(n,)).astype(np.float32).copy()
    n_val = [v.n_val[k] for k in range(v.n_par)]
    rows = int(np.prod(n_val))
    return dict(n_par=v.n_par, pid=[v.pid[k] for k in range(v.n_par)], n_val=n_val,
                val=[f(v.val[k], n_val[k]) for k in range(v.n_par)], n_ener=v.n_ener,
                elo=f(v.elo, v.n_ener), ehi=f(v.ehi, v.n_ener),
                spec=f(v.spec, rows * v.n_ener).reshape(rows, v.n_ener))


def rradtable_synthetic(seed=7, n_spin=21, n_rad=50, n_g=20):
    _check(lib.rx_rradtable_synthetic(seed, n_spin, n_rad, n_g))


def rradtable_view():
    """Copy of the resident returning-radiation table (spin-major arrays)."""
    v = _RradView()
    _check(lib.rx_rradtable_get_view(ctypes.byref(v)))
    ns, nr, ng = v.n_spin, v.n_rad, v.n_g
    f = lambda ptr, n: np.ctypeslib.as_array(ptr, shape=(n,)).copy()
    return dict(n_spin=ns, n_rad=nr, n_g=ng, spin=f(v.spin, ns), rlo=f(v.rlo, ns * nr), rhi=f(v.rhi, ns * nr),
                tf_r=f(v.tf_r, ns * nr * nr), gmin=f(v.gmin, ns * nr * nr), gmax=f(v.gmax, ns * nr * nr),
                frac_g=f(v.frac_g, ns * nr * nr * ng))


def tables_synthetic(seed=1, n_a=25, n_mu0=22, n_r=100, n_g=20):
    _check(lib.rx_tables_synthetic(seed, n_a, n_mu0, n_r, n_g))


def lptable_synthetic(seed=3, n_a=20, n_h=100, n_r=100):
    _check(lib.rx_lptable_synthetic(seed, n_a, n_h, n_r))


def set_conv_grid(n=4096, emin=0.00035, emax=2000.0):
    _check(lib.rx_set_conv_grid(n, emin, emax))


def tables_view():
    """Copies of the resident tables as numpy arrays (rel, lp dicts)."""
    rv, lv = _RelView(), _LpView()
    _check(lib.rx_tables_view(ctypes.byref(rv), ctypes.byref(lv)))
    nn = rv.n_a * rv.n_mu0

    def f(ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(np.float32).copy()

    rel = dict(n_a=rv.n_a, n_mu0=rv.n_mu0, n_r=rv.n_r, n_g=rv.n_g,
               a=f(rv.a, rv.n_a), mu0=f(rv.mu0, rv.n_mu0),
               r=f(rv.r, nn * rv.n_r), gmin=f(rv.gmin, nn * rv.n_r), gmax=f(rv.gmax, nn * rv.n_r),
               trff1=f(rv.trff1, nn * rv.n_r * rv.n_g), trff2=f(rv.trff2, nn * rv.n_r * rv.n_g),
               cosne1=f(rv.cosne1, nn * rv.n_r * rv.n_g), cosne2=f(rv.cosne2, nn * rv.n_r * rv.n_g))
    n3 = lv.n_a * lv.n_h * lv.n_r
    lp = dict(n_a=lv.n_a, n_h=lv.n_h, n_r=lv.n_r, a=f(lv.a, lv.n_a), h=f(lv.h, lv.n_a * lv.n_h),
              rad=f(lv.rad, lv.n_a * lv.n_r), intens=f(lv.intens, n3), del_=f(lv.del_, n3),
              del_inc=f(lv.del_inc, n3))
    return rel, lp


def reflset_create(refl_energy, refl):
    """Resident reflection set; refl has shape (n_incl, n_refl)."""
    e = _arr(refl_energy)
    r = _arr(refl)
    n_incl, n_refl = r.shape
    sid = lib.rx_reflset_create(_dp(e), n_refl, _dp(r.reshape(-1)), n_incl)
    if sid < 0:
        raise RxError(-sid, lib.rx_last_error().decode())
    return sid


class Batch:
    """Device-resident batch of requests (rx_batch_*)."""

    def __init__(self, requests):
        # requests: list of dicts with model, energy, params, [flux_in], [refl_set]
        self._keep = []
        arr = (_Request * len(requests))()
        self.n_flux = []
        for i, r in enumerate(requests):
            e = _arr(r["energy"])
            p = _arr(r["params"])
            self._keep += [e, p]
            arr[i].model = r["model"]
            arr[i].n_flux = e.size - 1
            arr[i].energy = _dp(e)
            arr[i].parameter = _dp(p)
            if r.get("flux_in") is not None:
                f = _arr(r["flux_in"])
                self._keep.append(f)
                arr[i].flux_in = _dp(f)
            else:
                arr[i].flux_in = None
            arr[i].refl_set = r.get("refl_set", -1)
            self.n_flux.append(e.size - 1)
        h = lib.rx_batch_create(arr, len(requests))
        if not h:
            raise RxError(lib.rx_last_status(), lib.rx_last_error().decode())
        self.h = h
        self.n = len(requests)
        self._keep = []

    def run(self, precision=FP64):
        _check(lib.rx_batch_run(self.h, precision))

    def sync(self):
        _check(lib.rx_batch_sync(self.h))

    def output(self, i):
        out = np.zeros(self.n_flux[i])
        st = lib.rx_batch_output(self.h, i, _dp(out))
        return out, st

    def outputs(self):
        return [self.output(i) for i in range(self.n)]

    def dist(self, i, n_incl):
        d = np.zeros(n_incl)
        _check(lib.rx_batch_dist(self.h, i, _dp(d), n_incl))
        return d

    def zones(self, i, n_zones):
        """(n_zones, 8): rmean, del_emit, shift, ncf, Ecut index, Ecut factor,
        logXi index, logXi factor"""
        o = np.zeros(8 * n_zones)
        _check(lib.rx_batch_zones(self.h, i, _dp(o), n_zones))
        return o.reshape(n_zones, 8)

    def reflfrac(self, i):
        o = np.zeros(5)
        _check(lib.rx_batch_reflfrac(self.h, i, _dp(o)))
        return o

    def set_shape(self, seg_bins, waves_per_cu):
        """bins per work item and persistent wavefronts per CU (0: automatic)"""
        _check(lib.rx_batch_set_shape(self.h, seg_bins, waves_per_cu))

    def autotune(self, precision):
        """returns (seg_bins, waves_per_cu)"""
        r = lib.rx_batch_autotune(self.h, precision)
        if r <= 0:
            raise RxError(-r, lib.rx_last_error().decode())
        return r // 256, r % 256

    def kernel_ms(self):
        """per kernel ms of the last profiled run: setup, integrate, fixup, finalize, conv, xill, primary"""
        o = np.zeros(7)
        _check(lib.rx_batch_kernel_ms(self.h, _dp(o)))
        return o

    def graph_replays(self):
        return int(lib.rx_batch_graph_replays(self.h))

    def counters(self):
        """work items, FP64 re-checks and re-check overflow of the last run"""
        o = (ctypes.c_int64 * 3)()
        _check(lib.rx_batch_counters(self.h, o))
        return dict(items=o[0], fp64_rechecks=o[1], recheck_overflow=o[2])

    def close(self):
        if getattr(self, "h", None):
            lib.rx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching serving loop (rx_server_*)."""

    def __init__(self, max_inflight=4096, hbm_budget=8 << 30, n_streams=4, precision=FP32):
        h = lib.rx_server_start(max_inflight, hbm_budget, n_streams, precision)
        if not h:
            raise RxError(lib.rx_last_status(), lib.rx_last_error().decode())
        self.h = h
        self._nflux = {}

    def submit(self, model, energy, params, flux_in=None, refl_set=-1):
        e = _arr(energy)
        p = _arr(params)
        r = _Request()
        r.model = model
        r.n_flux = e.size - 1
        r.energy = _dp(e)
        r.parameter = _dp(p)
        f = None
        if flux_in is not None:
            f = _arr(flux_in)
            r.flux_in = _dp(f)
        else:
            r.flux_in = None
        r.refl_set = refl_set
        t = lib.rx_server_submit(self.h, ctypes.byref(r))
        if t < 0:
            raise RxError(-t, "submit failed")
        self._nflux[t] = e.size - 1
        return t

    def replay(self, requests, arrival_s):
        """Submits requests[i] (dicts as for Batch) at arrival_s[i] seconds
        after the call, paced by the library's native loop; returns the
        tickets once the last one is submitted."""
        n = len(requests)
        arr = (_Request * n)()
        keep = []
        for i, r in enumerate(requests):
            e = _arr(r["energy"])
            p = _arr(r["params"])
            keep += [e, p]
            arr[i].model = r["model"]
            arr[i].n_flux = e.size - 1
            arr[i].energy = _dp(e)
            arr[i].parameter = _dp(p)
            if r.get("flux_in") is not None:
                f = _arr(r["flux_in"])
                keep.append(f)
                arr[i].flux_in = _dp(f)
            else:
                arr[i].flux_in = None
            arr[i].refl_set = r.get("refl_set", -1)
        t = np.ascontiguousarray(np.asarray(arrival_s, dtype=np.float64))
        tickets = (ctypes.c_int64 * n)()
        _check(lib.rx_server_replay(self.h, arr, n, _dp(t), tickets))
        out = list(tickets)
        for i, tk in enumerate(out):
            if tk < 0:
                raise RxError(-tk, "submit failed")
            self._nflux[tk] = arr[i].n_flux
        return out

    def poll(self, ticket):
        """None while pending, else (flux, status, latency_ms)."""
        out = np.zeros(self._nflux[ticket])
        st = ctypes.c_int(0)
        lat = np.zeros(1)
        r = lib.rx_server_poll(self.h, ticket, _dp(out), ctypes.byref(st), _dp(lat))
        if r == 0:
            return None
        if r < 0:
            raise RxError(r, "unknown ticket")
        del self._nflux[ticket]
        return out, st.value, float(lat[0])

    def drain(self):
        _check(lib.rx_server_drain(self.h))

    def stats(self):
        o = (ctypes.c_int64 * 3)()
        _check(lib.rx_server_stats(self.h, o))
        return dict(completed=o[0], batches=o[1], max_concurrent_batches=o[2])

    def stop(self):
        if getattr(self, "h", None):
            lib.rx_server_stop(self.h)
            self.h = None

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
