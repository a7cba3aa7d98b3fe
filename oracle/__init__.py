"""Python binding of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline; never
by the product library or its Python mirror.

The oracle restates relxill 2.4's relline / relline_lp / relconv algorithm in
plain C (see relxill_oracle.c for the per-function reference citations).
"""
import ctypes
import os

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_DIR, "liboracle.so")

MOD_RELLINE = 1
MOD_RELLINELP = 2
MOD_RELCONV = 11
MOD_RELCONVLP = 12
MOD_RELXILL = -1
MOD_RELXILLLP = -2
MOD_RELXILL_CONV = -1001
MOD_RELXILLCP = -1010
MOD_RELXILLLPCP = -1011
MOD_RELXILLNS = -1012
MOD_XILLVER = 0
MOD_XILLVERCP = 100
MOD_XILLVERNS = -101

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.c_int


class RelTable(ctypes.Structure):
    _fields_ = [("n_a", _I), ("n_mu0", _I), ("n_r", _I), ("n_g", _I), ("a", _F), ("mu0", _F),
                ("r", _F), ("gmin", _F), ("gmax", _F), ("trff1", _F), ("trff2", _F),
                ("cosne1", _F), ("cosne2", _F)]


class RradTable(ctypes.Structure):
    _fields_ = [("n_spin", _I), ("n_rad", _I), ("n_g", _I)] + \
               [(k, _D) for k in ("spin", "rlo", "rhi", "tf_r", "gmin", "gmax", "frac_g")]


class LpTable(ctypes.Structure):
    _fields_ = [("n_a", _I), ("n_h", _I), ("n_r", _I), ("a", _F), ("h", _F), ("rad", _F),
                ("intens", _F), ("del_", _F), ("del_inc", _F), ("rrad", ctypes.POINTER(RradTable))]


class RelParam(ctypes.Structure):
    _fields_ = [("model_type", _I), ("emis_type", _I), ("a", ctypes.c_double),
                ("incl", ctypes.c_double), ("emis1", ctypes.c_double), ("emis2", ctypes.c_double),
                ("rbr", ctypes.c_double), ("rin", ctypes.c_double), ("rout", ctypes.c_double),
                ("lineE", ctypes.c_double), ("z", ctypes.c_double), ("height", ctypes.c_double),
                ("htop", ctypes.c_double), ("gamma", ctypes.c_double), ("beta", ctypes.c_double),
                ("limb", _I), ("num_zones", _I), ("return_rad", _I), ("do_renorm", _I),
                ("rr_corr_flux", _D), ("rr_corr_gshift", _D), ("rr_corr_rgrid", _D), ("rr_corr_n", _I),
                ("ion_grad_type", _I), ("ion_grad_index", ctypes.c_double), ("lxi0", ctypes.c_double)]


class Stats(ctypes.Structure):
    _fields_ = [("relb_func_calls", ctypes.c_int64), ("romberg_calls", ctypes.c_int64),
                ("romberg_levels", ctypes.c_int64), ("edge_calls", ctypes.c_int64),
                ("midpoint_calls", ctypes.c_int64), ("bin_integrals", ctypes.c_int64)]


class SysPar(ctypes.Structure):
    _fields_ = [("nr", _I), ("ng", _I), ("re", _D), ("gmin", _D), ("gmax", _D), ("gstar", _D),
                ("d_gstar", _D), ("trff", _D), ("cosne", _D), ("emis", _D), ("del_emit", _D),
                ("del_inc", _D), ("limb_law", _I),
                ("refl", ctypes.c_double * 5)]


class XillTable(ctypes.Structure):
    _fields_ = [("n_par", _I), ("pid", _I * 6), ("n_val", _I * 6), ("val", _F * 6), ("n_ener", _I),
                ("elo", _F), ("ehi", _F), ("spec", _F)]


class XillParam(ctypes.Structure):
    _fields_ = [("gam", ctypes.c_double), ("afe", ctypes.c_double), ("lxi", ctypes.c_double),
                ("ect", ctypes.c_double), ("dens", ctypes.c_double), ("incl", ctypes.c_double),
                ("refl_frac", ctypes.c_double), ("boost", _I), ("prim", _I), ("ktbb", ctypes.c_double)]


class Job(ctypes.Structure):
    _fields_ = [("model_type", _I), ("parameter", _D), ("energy", _D), ("flux", _D),
                ("n_flux", _I), ("status", _I)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "orc_kerr_rms": (ctypes.c_double, [ctypes.c_double]),
        "orc_kerr_rplus": (ctypes.c_double, [ctypes.c_double]),
        "orc_ut_disk": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
        "orc_gi_potential_lp": (ctypes.c_double, [ctypes.c_double] * 5),
        "orc_lp_emissivity_newton": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
        "orc_get_log_grid": (None, [_D, _I, ctypes.c_double, ctypes.c_double]),
        "orc_get_fine_radial_grid": (None, [ctypes.c_double, ctypes.c_double, _D, _I]),
        "orc_binary_search": (_I, [_D, _I, ctypes.c_double]),
        "orc_inv_binary_search": (_I, [_D, _I, ctypes.c_double]),
        "orc_interp_lin_1d": (ctypes.c_double, [ctypes.c_double] * 3),
        "orc_interp_lin_2d": (ctypes.c_double, [ctypes.c_double] * 6),
        "orc_interp_lin_2d_float": (ctypes.c_double, [ctypes.c_double] * 2 + [ctypes.c_float] * 4),
        "orc_rebin_spectrum": (None, [_D, _D, _I, _D, _D, _I]),
        "orc_rebin_mean_flux": (None, [_D, _D, _I, _D, _D, _I]),
        "orc_get_rzone_grid": (_I, [ctypes.c_double, ctypes.c_double, _I, ctypes.c_double, _D]),
        "orc_get_rel_params": (_I, [_I, _D, _I, ctypes.POINTER(RelParam)]),
        "orc_syspar_new": (ctypes.POINTER(SysPar), [_I, _I]),
        "orc_syspar_free": (None, [ctypes.POINTER(SysPar)]),
        "orc_get_system_parameters": (ctypes.POINTER(SysPar),
                                      [ctypes.POINTER(RelTable), ctypes.POINTER(LpTable),
                                       ctypes.POINTER(RelParam), _I, ctypes.POINTER(_I)]),
        "orc_relbase": (_I, [ctypes.POINTER(RelTable), ctypes.POINTER(LpTable),
                             ctypes.POINTER(RelParam), _D, _I, _I, _D, _D, _I,
                             ctypes.POINTER(Stats)]),
        "orc_fft": (None, [_D, _D, _I, _I]),
        "orc_fft_conv_spectrum": (None, [_D, _D, _D, _D, _I]),
        "orc_fft_norm_factor": (ctypes.c_double, [_D, _D, _D, _D, _I]),
        "orc_conv_normalized": (None, [_D, _D, _D, _D, _I]),
        "orc_relconv_kernel": (_I, [ctypes.POINTER(RelTable), ctypes.POINTER(LpTable),
                                    ctypes.POINTER(RelParam), _D, _I, _D, _D, _I, _I,
                                    ctypes.POINTER(Stats)]),
        "orc_relxill_conv": (_I, [ctypes.POINTER(RelTable), ctypes.POINTER(RelParam), _D, _I,
                                  _D, _D, _I, _I, _D, _D, _I, _I, _D, ctypes.POINTER(Stats)]),
        "orc_xill_indices": (None, [ctypes.POINTER(XillTable), ctypes.POINTER(XillParam), ctypes.POINTER(_I)]),
        "orc_xill_factors": (None, [ctypes.POINTER(XillTable), ctypes.POINTER(XillParam), ctypes.POINTER(_I), _D]),
        "orc_xill_renorm": (None, [_F, _I, ctypes.c_double, ctypes.c_double]),
        "orc_xill_spectra": (None, [ctypes.POINTER(XillTable), ctypes.POINTER(XillParam), _D, _D]),
        "orc_xill_angdep": (None, [_D, _D, _D, _I, _I]),
        "orc_rrad_gshift_boost": (ctypes.c_double, [ctypes.c_double] * 3),
        "orc_xill_corr_factors": (None, [ctypes.POINTER(XillTable), ctypes.POINTER(XillParam), _D, _D]),
        "orc_spec_cutoffpl": (None, [_D, _D, _I, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "orc_norm_wrt_xillver": (ctypes.c_double, [_D, _D, _I, ctypes.POINTER(_I)]),
        "orc_primary_norm_factor": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
        "orc_prim_norm_factor": (ctypes.c_double, [ctypes.POINTER(XillParam), ctypes.c_double]),
        "orc_spec_blackbody": (None, [_D, _D, _I, ctypes.c_double]),
        "orc_xill_spectrum_incl": (None, [ctypes.POINTER(XillTable), ctypes.POINTER(XillParam), _D, _D]),
        "orc_xillver_model": (_I, [_I, ctypes.POINTER(XillTable), _D, _D, _D, _I]),
        "orc_prim_spectrum": (None, [_D, _D, _I, ctypes.POINTER(XillParam), ctypes.c_double, ctypes.c_double]),
        "orc_mcd_shape": (ctypes.c_double, [ctypes.c_double]),
        "orc_nthcomp_solve": (_I, [ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D]),
        "orc_nthcomp_photar": (None, [_D, _D, _I, _D, _I, ctypes.c_double, _D]),
        "orc_spec_nthcomp": (None, [_D, _D, _I, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "orc_get_relxill_params": (_I, [_I, _D, _I, ctypes.POINTER(RelParam), ctypes.POINTER(XillParam)]),
        "orc_energy_shift_source_obs": (ctypes.c_double, [ctypes.POINTER(RelParam)]),
        "orc_relxill_zones": (_I, [_I, ctypes.POINTER(RelTable), ctypes.POINTER(LpTable), _D, _D]),
        "orc_rrad_emis": (_I, [ctypes.POINTER(RradTable), ctypes.POINTER(RelParam), _D, _D, _I, _D]),
        "orc_relxill_model": (_I, [_I, ctypes.POINTER(RelTable), ctypes.POINTER(LpTable), ctypes.POINTER(XillTable),
                                   _D, _D, _D, _I, _D, _I, _I, _D, ctypes.POINTER(Stats)]),
        "orc_relxill": (_I, [ctypes.POINTER(RelTable), ctypes.POINTER(XillTable), _D, _D, _D, _I, _D, _I, _I,
                             _D, ctypes.POINTER(Stats)]),
        "orc_eval_model": (_I, [_I, ctypes.POINTER(RelTable), ctypes.POINTER(LpTable), _D, _D,
                                _D, _I, _D, _I, _I, ctypes.POINTER(Stats)]),
        "orc_eval_batch_mt": (_I, [ctypes.POINTER(RelTable), ctypes.POINTER(LpTable),
                                   ctypes.POINTER(Job), _I, _D, _I, _I]),
        "orc_eval_batch_mt_stats": (_I, [ctypes.POINTER(RelTable), ctypes.POINTER(LpTable),
                                         ctypes.POINTER(Job), _I, _D, _I, _I, ctypes.POINTER(Stats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


lib = _load()


def _arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _dp(a):
    return a.ctypes.data_as(_D)


def _fp(a):
    return a.ctypes.data_as(_F)


class Tables:
    """Oracle view of a relline table (and optional lamp-post table) given as
    dicts of numpy arrays in the layout of relxill_amd.tables_view()."""

    def __init__(self, rel, lp=None, rrad=None):
        self._keep = []

        def f32(k, d):
            a = np.ascontiguousarray(d[k], dtype=np.float32)
            self._keep.append(a)
            return a
        self.rel = RelTable(rel["n_a"], rel["n_mu0"], rel["n_r"], rel["n_g"],
                            *[_fp(f32(k, rel)) for k in ("a", "mu0", "r", "gmin", "gmax",
                                                         "trff1", "trff2", "cosne1", "cosne2")])
        self.lp = None
        self.rr = None
        if lp is not None:
            self.lp = LpTable(lp["n_a"], lp["n_h"], lp["n_r"],
                              *[_fp(f32(k, lp)) for k in ("a", "h", "rad", "intens", "del_",
                                                          "del_inc")])
            if rrad is not None:
                def f64(k):
                    a = np.ascontiguousarray(rrad[k], dtype=np.float64)
                    self._keep.append(a)
                    return _dp(a)
                self.rr = RradTable(rrad["n_spin"], rrad["n_rad"], rrad["n_g"],
                                    *[f64(k) for k in ("spin", "rlo", "rhi", "tf_r", "gmin", "gmax", "frac_g")])
                self.lp.rrad = ctypes.pointer(self.rr)
        self.rel_arrays = rel
        self.lp_arrays = lp

    @property
    def lp_ptr(self):
        return ctypes.byref(self.lp) if self.lp is not None else None


def rel_params(model, params, phys_norm=False):
    p = RelParam()
    st = lib.orc_get_rel_params(model, _dp(_arr(params)), int(phys_norm), ctypes.byref(p))
    return p, st


def eval_model(model, tables, params, energy, flux_in=None, conv_grid=None, phys_norm=False,
               stats=False):
    """Model-level evaluation (what lmod<name> computes); returns (flux, status[, stats])."""
    e = _arr(energy)
    n = e.size - 1
    f = _arr(flux_in).copy() if flux_in is not None else np.zeros(n)
    st_obj = Stats()
    cg = _arr(conv_grid) if conv_grid is not None else None
    st = lib.orc_eval_model(model, ctypes.byref(tables.rel), tables.lp_ptr, _dp(_arr(params)),
                            _dp(e), _dp(f), n, _dp(cg) if cg is not None else None,
                            (cg.size - 1) if cg is not None else 0, int(phys_norm),
                            ctypes.byref(st_obj))
    if stats:
        return f, st, {k: getattr(st_obj, k) for k, _ in Stats._fields_}
    return f, st


def relbase(tables, relparam, ener, nr_fine=1000, n_incl=0, stats=False):
    e = _arr(ener)
    n = e.size - 1
    flux = np.zeros(n * max(1, relparam.num_zones))
    dist = np.zeros(max(1, n_incl) * max(1, relparam.num_zones))
    so = Stats()
    st = lib.orc_relbase(ctypes.byref(tables.rel), tables.lp_ptr, ctypes.byref(relparam), _dp(e),
                         n, nr_fine, _dp(flux), _dp(dist) if n_incl > 0 else None, n_incl,
                         ctypes.byref(so))
    out = (flux, dist[:n_incl] if n_incl > 0 else None, st)
    if stats:
        return out + ({k: getattr(so, k) for k, _ in Stats._fields_},)
    return out


def system_parameters(tables, relparam, nr_fine=1000):
    """(re, gmin, gmax, trff[nr,ng,2], cosne, emis, del_emit, del_inc, refl5), status"""
    st = ctypes.c_int(0)
    sp = lib.orc_get_system_parameters(ctypes.byref(tables.rel), tables.lp_ptr,
                                       ctypes.byref(relparam), nr_fine, ctypes.byref(st))
    if not sp:
        return None, st.value
    s = sp.contents
    nr, ng = s.nr, s.ng
    g = lambda p, n: np.ctypeslib.as_array(p, shape=(n,)).copy()
    out = dict(re=g(s.re, nr), gmin=g(s.gmin, nr), gmax=g(s.gmax, nr),
               trff=g(s.trff, nr * ng * 2).reshape(nr, ng, 2),
               cosne=g(s.cosne, nr * ng * 2).reshape(nr, ng, 2), emis=g(s.emis, nr),
               del_emit=g(s.del_emit, nr), del_inc=g(s.del_inc, nr), refl=np.array(s.refl[:]),
               gstar=g(s.gstar, ng), d_gstar=g(s.d_gstar, ng))
    lib.orc_syspar_free(sp)
    return out, st.value


def rebin(ener, ener0, flu0):
    e = _arr(ener)
    e0 = _arr(ener0)
    f0 = _arr(flu0)
    out = np.zeros(e.size - 1)
    lib.orc_rebin_spectrum(_dp(e), _dp(out), e.size - 1, _dp(e0), _dp(f0), e0.size - 1)
    return out


def rebin_mean_flux(xn, x0, y0):
    xn = _arr(xn)
    x0 = _arr(x0)
    y0 = _arr(y0)
    out = np.zeros(xn.size - 1)
    lib.orc_rebin_mean_flux(_dp(xn), _dp(out), xn.size - 1, _dp(x0), _dp(y0), x0.size - 1)
    return out


def log_grid(n_edges, emin, emax):
    out = np.zeros(n_edges)
    lib.orc_get_log_grid(_dp(out), n_edges, emin, emax)
    return out


def fine_radial_grid(rin, rout, nr):
    out = np.zeros(nr)
    lib.orc_get_fine_radial_grid(rin, rout, _dp(out), nr)
    return out


def fft(x, inverse=False):
    x = np.asarray(x, dtype=np.complex128)
    re = np.ascontiguousarray(x.real.copy())
    im = np.ascontiguousarray(x.imag.copy())
    lib.orc_fft(_dp(re), _dp(im), x.size, int(inverse))
    return re + 1j * im


def fft_conv(ener, fxill, frel, normalized=True):
    e = _arr(ener)
    n = e.size - 1
    out = np.zeros(n)
    fn = lib.orc_conv_normalized if normalized else lib.orc_fft_conv_spectrum
    fn(_dp(e), _dp(_arr(fxill)), _dp(_arr(frel)), _dp(out), n)
    return out


def rzone_grid(rmin, rmax, nzones, h):
    out = np.zeros(nzones + 1)
    lib.orc_get_rzone_grid(rmin, rmax, nzones, h, _dp(out))
    return out


def relxill_conv(tables, params, conv_grid, refl_ener, refl, out_ener, nr_fine=1000,
                 phys_norm=False):
    p, st = rel_params(MOD_RELXILL_CONV, params, phys_norm)
    if st:
        return None, None, st
    cg = _arr(conv_grid)
    r = _arr(refl)
    n_incl, n_refl = r.shape
    oe = _arr(out_ener)
    out = np.zeros(oe.size - 1)
    dist = np.zeros(n_incl)
    st = lib.orc_relxill_conv(ctypes.byref(tables.rel), ctypes.byref(p), _dp(cg), cg.size - 1,
                              _dp(_arr(refl_ener)), _dp(r.reshape(-1)), n_refl, n_incl, _dp(oe),
                              _dp(out), oe.size - 1, nr_fine, _dp(dist), None)
    return out, dist, st


def eval_batch_mt(tables, jobs, n_threads, conv_grid=None, stats=False):
    """jobs: list of (model, params, energy, flux_in or None); returns (fluxes,
    statuses[, work counters summed over the jobs])."""
    arr = (Job * len(jobs))()
    keep = []
    outs = []
    for i, (m, p, e, f) in enumerate(jobs):
        pa = _arr(p)
        ea = _arr(e)
        fa = _arr(f).copy() if f is not None else np.zeros(ea.size - 1)
        keep += [pa, ea]
        outs.append(fa)
        arr[i] = Job(m, _dp(pa), _dp(ea), _dp(fa), ea.size - 1, 0)
    cg = _arr(conv_grid) if conv_grid is not None else None
    so = Stats()
    lib.orc_eval_batch_mt_stats(ctypes.byref(tables.rel), tables.lp_ptr, arr, len(jobs),
                                _dp(cg) if cg is not None else None,
                                (cg.size - 1) if cg is not None else 0, n_threads,
                                ctypes.byref(so) if stats else None)
    if stats:
        return outs, [arr[i].status for i in range(len(jobs))], {k: getattr(so, k) for k, _ in Stats._fields_}
    return outs, [arr[i].status for i in range(len(jobs))]


class Xill:
    """Oracle view of a xillver table given as the dict of
    relxill_amd.xilltable_view() (rows as stored in the file)."""

    def __init__(self, x):
        self._keep = []

        def f32(a):
            a = np.ascontiguousarray(a, dtype=np.float32)
            self._keep.append(a)
            return a
        t = XillTable()
        t.n_par = x["n_par"]
        for k in range(x["n_par"]):
            t.pid[k] = x["pid"][k]
            t.n_val[k] = x["n_val"][k]
            t.val[k] = _fp(f32(x["val"][k]))
        t.n_ener = x["n_ener"]
        t.elo = _fp(f32(x["elo"]))
        t.ehi = _fp(f32(x["ehi"]))
        t.spec = _fp(f32(np.asarray(x["spec"]).reshape(-1)))
        self.t = t
        self.arrays = x

    @property
    def n_incl(self):
        return self.arrays["n_val"][-1]


def xill_param(gam, afe, lxi, ect, dens=15.0, incl=30.0, refl_frac=0.0, prim=0, ktbb=0.0):
    """xillver parameters; prim 1: the nthcomp primary (ect is kTe), 2: the
    blackbody (ktbb)"""
    return XillParam(gam, afe, lxi, ect, dens, incl, refl_frac, 0, prim, ktbb)


def prim_norm_factor(gam, ect, prim=0, ktbb=0.0):
    """1 / calcNormWrtXillverTableSpec of the primary (cut-off power law,
    nthcomp with prim=1, the blackbody of ktbb with prim=2) on the 500-bin grid"""
    return lib.orc_prim_norm_factor(ctypes.byref(xill_param(gam, 1.0, 0.0, ect, prim=prim, ktbb=ktbb)), ect)


def xill_spectrum_incl(xt, xp):
    """(ener[n+1], flu[n]) of the table for the xillver models (the
    inclination interpolated)"""
    n = xt.t.n_ener
    e = np.zeros(n + 1)
    flu = np.zeros(n)
    lib.orc_xill_spectrum_incl(ctypes.byref(xt.t), ctypes.byref(xp), _dp(e), _dp(flu))
    return e, flu


def xillver_model(model, xt, params, energy):
    """xillver / xillverCp / xillverNS on the caller grid; (flux, status)"""
    e = _arr(energy)
    out = np.zeros(e.size - 1)
    st = lib.orc_xillver_model(model, ctypes.byref(xt.t), _dp(_arr(params)), _dp(e), _dp(out), e.size - 1)
    return out, st


def blackbody(ener, ktbb):
    e = _arr(ener)
    out = np.zeros(e.size - 1)
    lib.orc_spec_blackbody(_dp(out), _dp(e), e.size - 1, ktbb)
    return out


def xill_corner(xt, xp):
    """(indices, factors) of get_xilltab_indices_for_paramvals / interp_xill_table"""
    ind = (_I * 6)()
    fac = np.zeros(6)
    lib.orc_xill_indices(ctypes.byref(xt.t), ctypes.byref(xp), ind)
    lib.orc_xill_factors(ctypes.byref(xt.t), ctypes.byref(xp), ind, _dp(fac))
    return np.array(ind[:xt.t.n_par]), fac[:xt.t.n_par]


def xill_spectra(xt, xp):
    """(ener[n+1], flu[n_incl, n]) of the table for the relxill model"""
    n = xt.t.n_ener
    e = np.zeros(n + 1)
    flu = np.zeros(xt.n_incl * n)
    lib.orc_xill_spectra(ctypes.byref(xt.t), ctypes.byref(xp), _dp(e), _dp(flu))
    return e, flu.reshape(xt.n_incl, n)


def xill_corr_factors(xt, xp):
    """(corr_flux, corr_gshift) of get_xillver_fluxcorrection_factors"""
    a, b = np.zeros(1), np.zeros(1)
    lib.orc_xill_corr_factors(ctypes.byref(xt.t), ctypes.byref(xp), _dp(a), _dp(b))
    return float(a[0]), float(b[0])


def cutoffpl(ener, ect, gam, shift=1.0):
    e = _arr(ener)
    out = np.zeros(e.size - 1)
    lib.orc_spec_cutoffpl(_dp(out), _dp(e), e.size - 1, ect, gam, shift)
    return out


def primary_norm_factor(gam, ect):
    return lib.orc_primary_norm_factor(gam, ect)


def nthcomp_solve(kte, gam, ktbb=0.05):
    """the Kompaneets solution for a disc seed: (x [m_e c^2], E F_E)"""
    x = np.zeros(901)
    spt = np.zeros(900)
    n = lib.orc_nthcomp_solve(ktbb / 511.0, kte / 511.0, gam, _dp(x), _dp(spt))
    return x[:n + 1].copy(), spt[:n].copy()


def nthcomp_photar(x, spt, ener, z=0.0):
    e = _arr(ener)
    xs, ss = _arr(x), _arr(spt)
    out = np.zeros(e.size - 1)
    lib.orc_nthcomp_photar(_dp(xs), _dp(ss), ss.size, _dp(e), e.size - 1, z, _dp(out))
    return out


def nthcomp(ener, kte, gam, shift=1.0):
    """spec_nthcomp: photons per bin, E F_E(1 keV in the source) = 1"""
    e = _arr(ener)
    out = np.zeros(e.size - 1)
    lib.orc_spec_nthcomp(_dp(out), _dp(e), e.size - 1, kte, gam, shift)
    return out


def relxill(tables, xt, params, energy, conv_grid=None, phys_norm=False):
    """relxill on the caller grid; returns (flux, dist, status)."""
    e = _arr(energy)
    n = e.size - 1
    out = np.zeros(n)
    dist = np.zeros(xt.n_incl)
    cg = _arr(conv_grid) if conv_grid is not None else None
    st = lib.orc_relxill(ctypes.byref(tables.rel), ctypes.byref(xt.t), _dp(_arr(params)), _dp(e), _dp(out), n,
                         _dp(cg) if cg is not None else None, (cg.size - 1) if cg is not None else 0,
                         int(phys_norm), _dp(dist), None)
    return out, dist, st


def relxill_model(model, tables, xt, params, energy, conv_grid=None, phys_norm=False):
    """relxill (MOD_RELXILL), relxilllp (MOD_RELXILLLP) or their nthcomp
    forms (MOD_RELXILLCP, MOD_RELXILLLPCP) on the caller grid;
    returns (flux, dist[num_zones, n_incl], status)."""
    e = _arr(energy)
    n = e.size - 1
    out = np.zeros(n)
    nz = n_zones(model, params)
    dist = np.zeros(nz * xt.n_incl)
    cg = _arr(conv_grid) if conv_grid is not None else None
    st = lib.orc_relxill_model(model, ctypes.byref(tables.rel), tables.lp_ptr, ctypes.byref(xt.t),
                               _dp(_arr(params)), _dp(e), _dp(out), n, _dp(cg) if cg is not None else None,
                               (cg.size - 1) if cg is not None else 0, int(phys_norm), _dp(dist), None)
    return out, dist.reshape(nz, xt.n_incl), st


def n_zones(model, params):
    """radial zones of a relxill model: 10 for the lamp post, 25 for
    relxilllpCp with the power-law ionisation gradient (iongrad_type 1), else 1"""
    if model == MOD_RELXILLLPCP and round(params[14]) == 1:
        return 25
    return 10 if model in (MOD_RELXILLLP, MOD_RELXILLLPCP) else 1


def relxill_zones(model, tables, params):
    """per zone: r_lo, r_hi, del_emit(centre), energy shift source -> disk,
    Ecut seen by the zone, normalisation change, logXi; (array[nz, 7], status)"""
    nz = n_zones(model, params)
    out = np.zeros(7 * nz)
    st = lib.orc_relxill_zones(model, ctypes.byref(tables.rel), tables.lp_ptr, _dp(_arr(params)), _dp(out))
    return out.reshape(nz, 7), st


def energy_shift_source_obs(model, params):
    p, st = rel_params(model, params)
    return lib.orc_energy_shift_source_obs(ctypes.byref(p)), st


def rrad_emis(rrad, relparam, re, emis0):
    """returning-radiation emissivity on the fine grid re (descending); (array, status)"""
    t = Tables.__new__(Tables)
    keep = []

    def f64(k):
        a = np.ascontiguousarray(rrad[k], dtype=np.float64)
        keep.append(a)
        return _dp(a)
    rt = RradTable(rrad["n_spin"], rrad["n_rad"], rrad["n_g"],
                   *[f64(k) for k in ("spin", "rlo", "rhi", "tf_r", "gmin", "gmax", "frac_g")])
    re = _arr(re)
    e0 = _arr(emis0)
    out = np.zeros(re.size)
    st = lib.orc_rrad_emis(ctypes.byref(rt), ctypes.byref(relparam), _dp(re), _dp(e0), re.size, _dp(out))
    return out, st
