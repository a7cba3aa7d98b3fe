"""rx_eval_model_noalloc, the allocation-free entry point, against
rx_eval_model through the C ABI: the same outputs (only the FP64 order of the
per-chunk sums may differ) for the four line and convolution models in both
precisions, on two calls in a row whose second has the smaller grid (the
scratch array keeps the first call's contents); the input spectrum of a
convolution may be the output buffer; requests that need more work memory
than the array holds are refused with a clear error."""
import ctypes

import numpy as np
import pytest

import relxill_amd as rx
from tests.helpers import abs_err, draw_lp_params, log_grid, rel_err, tables

pytestmark = pytest.mark.gpu

_D = ctypes.POINTER(ctypes.c_double)
_H2D, _D2H = 1, 2


@pytest.fixture(scope="module")
def tabs():
    rel, lp, T = tables()
    assert rx.lib.rx_init(0) == 0, rx.lib.rx_last_error()
    return T


def _hip():
    """The HIP runtime the library runs on (the one already mapped into the
    process), for the caller-owned device buffers."""
    path = next(ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64" in ln)
    h = ctypes.CDLL(path)
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


class _Dev:
    """A device array of doubles, filled from `a`."""

    def __init__(self, hip, a):
        self.hip = hip
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.n = a.size
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), 8 * max(self.n, 1)) == 0
        assert hip.hipMemcpy(self.ptr, a.ctypes.data, 8 * self.n, _H2D) == 0

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, _D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc(model, params, e, flux_in, precision, in_place=False):
    """One call of the allocation-free entry point on the null stream; the
    output buffer starts out holding 7.0 in every bin (or the input
    spectrum, in place)."""
    hip = _hip()
    n = e.size - 1
    p = np.ascontiguousarray(params, dtype=np.float64)
    d_e = _Dev(hip, e)
    d_out = _Dev(hip, flux_in if (in_place and flux_in is not None) else np.full(n, 7.0))
    d_in = d_out if in_place else (_Dev(hip, flux_in) if flux_in is not None else None)
    try:
        st = rx.lib.rx_eval_model_noalloc(model, p.ctypes.data_as(_D), d_e.ptr, d_in.ptr if d_in else None,
                                          d_out.ptr, n, precision, None)
        return d_out.get(), st
    finally:
        for d in {id(x): x for x in (d_e, d_out, d_in) if x is not None}.values():
            d.free()


def _input(e):
    em = 0.5 * (e[1:] + e[:-1])
    return (em ** -2.0 * np.exp(-em / 300.0) + 0.02 * np.exp(-0.5 * ((em - 6.4) / 0.05) ** 2)) * np.diff(e)


CASES = [
    (rx.MODEL_RELLINE, [6.4, 3.0, 3.0, 15.0, 0.998, 30.0, -1.0, 400.0, 0.0, 0.0], (0.1, 10.0)),
    (rx.MODEL_RELLINELP, draw_lp_params(1, seed=12)[0], (0.1, 10.0)),
    (rx.MODEL_RELCONV, [3.0, 2.5, 12.0, -0.4, 60.0, 3.0, 400.0, 1.0], (0.1, 100.0)),
    (rx.MODEL_RELCONVLP, [8.0, 0.2, 0.7, 35.0, -1.0, 300.0, 0.0, 2.0, 0.0], (0.1, 100.0)),
]


@pytest.mark.parametrize("model,params,band", CASES)
@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_equals_rx_eval_model(tabs, model, params, band, precision):
    conv = model in (rx.MODEL_RELCONV, rx.MODEL_RELCONVLP)
    for n in (3000, 700):  # the second call is the smaller one
        e = log_grid(n, *band)
        f_in = _input(e) if conv else None
        ref, st_ref = rx.eval_model(model, params, e, flux_in=f_in, precision=precision)
        y, st = _noalloc(model, params, e, f_in, precision)
        err = abs_err(y, ref) if conv else rel_err(y, ref)
        print(f"noalloc model {model} prec {precision} n {n}: status {st} difference {err:.3e}", flush=True)
        assert st == 0 and st_ref == 0, rx.lib.rx_last_error()
        assert err < (1e-13 if conv else 1e-12)


def test_noalloc_in_place_convolution_and_errors(tabs):
    e = log_grid(2000, 0.1, 100.0)
    f_in = _input(e)
    p = CASES[2][1]
    ref, _ = rx.eval_model(rx.MODEL_RELCONV, p, e, flux_in=f_in, precision=rx.FP64)
    y, st = _noalloc(rx.MODEL_RELCONV, p, e, f_in, rx.FP64, in_place=True)
    assert st == 0 and abs_err(y, ref) < 1e-13
    # parameters the reference rejects: zero flux and the status
    bad = list(CASES[0][1])
    bad[6], bad[7] = 50.0, 40.0
    y, st = _noalloc(rx.MODEL_RELLINE, bad, log_grid(500, 0.1, 10.0), None, rx.FP64)
    assert st == rx.RX_ERR_PARAM and np.all(y == 0)
    # an input spectrum summing to zero (src/LocalModel.cpp:84-86)
    y, st = _noalloc(rx.MODEL_RELCONV, p, e, np.zeros(2000), rx.FP64)
    assert st == rx.RX_ERR_INPUT and np.all(y == 0)
    # more work memory than the scratch array holds: refused before any launch
    e_big = log_grid(600000, 0.1, 10.0)
    y, st = _noalloc(rx.MODEL_RELLINE, CASES[0][1], e_big, None, rx.FP64)
    assert st == rx.RX_ERR_ARG
    assert "RX_NOALLOC_SCRATCH_BYTES" in rx.lib.rx_last_error().decode()
    # relxill models are not offered here
    y, st = _noalloc(rx.MODEL_RELXILL, [3.0] * 13, e, None, rx.FP64)
    assert st == rx.RX_ERR_ARG
