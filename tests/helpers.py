"""Shared inputs and metrics for the parity tests.

Workloads follow the benchmark configurations: a 25 x 22 x 100 x 20 synthetic
relline table (seed 1), a 20 x 100 x 100 lamp-post table (seed 3), 3000 log
bins over 0.1-10 keV for relline, the reference convolution grid (4096 log bins
over 0.00035-2000 keV) for relconv.
"""
import numpy as np

import oracle
import relxill_amd as rx

RELLINE_CFG0 = [6.4, 3.0, 3.0, 15.0, 0.998, 30.0, -1.0, 400.0, 0.0, 0.0]


def log_grid(n_bins, emin, emax):
    """get_log_grid (src/relutility.c:399-405) with n_bins+1 edges."""
    return oracle.log_grid(n_bins + 1, emin, emax)


_TABLES = {}


def tables():
    """(rx tables view, oracle.Tables) of the resident default tables (the
    oracle's with the default returning-radiation table too)."""
    if "t" not in _TABLES:
        rx.tables_synthetic(1, 25, 22, 100, 20)
        rx.lptable_synthetic(3, 20, 100, 100)
        rx.rradtable_synthetic(7, 21, 50, 20)
        rel, lp = rx.tables_view()
        _TABLES["t"] = (rel, lp, oracle.Tables(rel, lp, rx.rradtable_view()))
    return _TABLES["t"]


def rel_err(y, yref, floor=1e-10):
    """max |y - y*|/|y*| over bins with |y*| > floor * max|y*| (the reference's
    comparison filter, test/refdata/test_refdata_relxill.sl:162-163)."""
    y = np.asarray(y, dtype=np.float64)
    yref = np.asarray(yref, dtype=np.float64)
    m = np.abs(yref).max()
    if m == 0:
        return float(np.abs(y).max())
    sel = np.abs(yref) > floor * m
    return float(np.max(np.abs(y[sel] - yref[sel]) / np.abs(yref[sel])))


def conv_err(y, yref, level=1e-5):
    """max |y - y*| / (|y*| + level max|y*|): the error relative to the bin,
    or to `level` times the maximum where the bin is smaller.  The relxill
    models' FP64 FFT convolutions (one per zone, each renormalised) carry a
    smooth absolute round-off of up to ~1e-13 of the spectrum's maximum that
    differs between two FFT implementations (6e-14 measured between the
    device and the oracle for ten-zone lamp-post requests, in every bin); a
    bin 1e-7 below the maximum sees it as 1e-6 relative.  With this measure a
    tolerance of 1e-8 means 1e-8 relative down to 1e-5 of the maximum and
    1e-13 of the maximum below that."""
    y = np.asarray(y, dtype=np.float64)
    yref = np.asarray(yref, dtype=np.float64)
    m = np.abs(yref).max()
    if m == 0:
        return float(np.abs(y).max())
    return float(np.max(np.abs(y - yref) / (np.abs(yref) + level * m)))


def abs_err(y, yref):
    """unfiltered max |y - y*| / max |y*|"""
    m = np.abs(yref).max()
    return float(np.abs(np.asarray(y) - np.asarray(yref)).max() / (m if m else 1.0))


def reflection_spectrum(n_bins=2999, emin=0.07, emax=1000.0, n_incl=10, seed=2):
    """Seeded synthetic reflection spectra (photons/bin) on a log grid: a cutoff
    power law (Gamma=2, Ecut=300 keV) plus Gaussian lines at 6.4 and 6.97 keV,
    with a smooth per-angle modulation.  Returns (edges, spectra[n_incl, n_bins])."""
    rng = np.random.default_rng(seed)
    e = log_grid(n_bins, emin, emax)
    em = 0.5 * (e[1:] + e[:-1])
    de = np.diff(e)
    base = em ** -2.0 * np.exp(-em / 300.0)
    lines = 0.02 * np.exp(-0.5 * ((em - 6.4) / 0.05) ** 2) + 0.008 * np.exp(-0.5 * ((em - 6.97) / 0.05) ** 2)
    out = np.empty((n_incl, n_bins))
    for k in range(n_incl):
        mu = 1.0 - (k + 0.5) / n_incl
        amp = 1.0 + 0.3 * (mu - 0.5) + 0.05 * rng.standard_normal()
        tilt = 0.05 * (mu - 0.5)
        out[k] = (base * em ** tilt + lines * (1.0 + 0.5 * mu)) * amp * de
    return e, out


def draw_lp_params(n, seed=4, lineE=6.4):
    """cfg2 priors: a U[-0.998,0.998], h U[3,100], incl U[5,85] deg, Rin U[ISCO,50]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = rng.uniform(-0.998, 0.998)
        h = rng.uniform(3.0, 100.0)
        incl = rng.uniform(5.0, 85.0)
        isco = oracle.lib.orc_kerr_rms(a)
        rin = rng.uniform(isco, 50.0)
        out.append([lineE, h, a, incl, rin, 400.0, 0.0, 0.0, 2.0, 0.0])
    return out


def draw_relconv_params(n, seed=5):
    """cfg3 priors with a broken power law q=3: relconv parameter arrays."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = rng.uniform(-0.998, 0.998)
        incl = rng.uniform(5.0, 85.0)
        isco = oracle.lib.orc_kerr_rms(a)
        rin = rng.uniform(isco, 50.0)
        out.append([3.0, 3.0, 15.0, a, incl, rin, 400.0, 0.0])
    return out


def flat_table(gmin, gmax, trff_sum=1.0, cosne=0.5):
    """The default table's a / mu0 / radius grids with a constant window
    [gmin, gmax] and a flat transfer function trff1 = trff2 = trff_sum/2."""
    rel, lp, T = tables()
    out = dict(rel)
    out["gmin"] = np.full_like(rel["gmin"], gmin)
    out["gmax"] = np.full_like(rel["gmax"], gmax)
    out["trff1"] = np.full_like(rel["trff1"], trff_sum / 2)
    out["trff2"] = np.full_like(rel["trff2"], trff_sum / 2)
    out["cosne1"] = np.full_like(rel["cosne1"], cosne)
    out["cosne2"] = np.full_like(rel["cosne2"], cosne)
    return out


def write_rxtb(path, rel):
    """Writes a relline table dict (tables_view layout) in the library's RXTB
    format: int32 magic 'RXTB', kind 1, n_a, n_mu0, n_r, n_g, then the float32
    arrays a, mu0, r, gmin, gmax, trff1, trff2, cosne1, cosne2."""
    hdr = np.array([0x42545852, 1, rel["n_a"], rel["n_mu0"], rel["n_r"], rel["n_g"]], dtype=np.int32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        for k in ("a", "mu0", "r", "gmin", "gmax", "trff1", "trff2", "cosne1", "cosne2"):
            f.write(np.ascontiguousarray(rel[k], dtype=np.float32).tobytes())


def sawtooth_table(c_rel=0.5):
    """Flat window [0.999, 1.001] with trff_k[g] = (1 + c g/(n_g-1))/2 for every
    node and radius: relb_func's swapped lerp (src/Relprofile.cpp:502-504) turns
    this into a sawtooth in g* with known per-bin integrals."""
    rel = flat_table(0.999, 1.001)
    ng = rel["n_g"]
    nn = rel["n_a"] * rel["n_mu0"] * rel["n_r"]
    saw = 0.5 * (1.0 + c_rel / (ng - 1) * np.arange(ng))
    rel["trff1"] = np.tile(saw, nn).astype(np.float32)
    rel["trff2"] = rel["trff1"].copy()
    return rel


def sawtooth_bins(e, rel, c_rel=0.5, swapped=True):
    """Closed-form photons per bin of relline on sawtooth_table() with the
    physical normalisation: F_j = 1/4 int E^3 f(g*) / sqrt(g*(1-g*)) dg* / E_mid
    (1/4 = sum of emis * pi r dr / 2 for a normalised emissivity), f the
    branch sum of the lerped transfer function, for the bins in 0.02 < g* <
    0.98 that contain no g* node.  Returns (indices, values)."""
    from scipy.integrate import quad
    ng = rel["n_g"]
    H = 5e-3
    step = (1 - 2 * H) / (ng - 1)
    c = c_rel / (ng - 1)
    gmin, gmax = float(rel["gmin"][0]), float(rel["gmax"][0])

    def f(g):
        q = (g - H) / step
        i = min(max(np.floor(q), 0), ng - 2)
        inte = q - i
        return 1 + c * (i + 1 - inte) if swapped else 1 + c * (i + inte)

    def integrand(g):
        E = gmin + g * (gmax - gmin)
        return E ** 3 * f(g) / np.sqrt(g * (1 - g))

    nodes = H + step * np.arange(ng)
    idx, val = [], []
    for j in range(len(e) - 1):
        glo = (e[j] - gmin) / (gmax - gmin)
        ghi = (e[j + 1] - gmin) / (gmax - gmin)
        if glo < 0.02 or ghi > 0.98 or np.any((nodes > glo) & (nodes < ghi)):
            continue
        idx.append(j)
        val.append(0.25 * quad(integrand, glo, ghi, epsabs=0, epsrel=1e-12)[0] / (0.5 * (e[j] + e[j + 1])))
    return np.array(idx), np.array(val)


def affine_table(alpha=1.0, beta=0.3, gamma=0.5):
    """Flat window with trff_1 = trff_2 = (alpha + beta a + gamma mu0)/2 at
    every node: the bilinear (a, mu0) interpolation (src/Relprofile.cpp:39-80,
    src/relutility.c:53-58) reproduces an affine field exactly, so the flux is
    pi/4 * (alpha + beta a_f + gamma mu0_f) with a_f, mu0_f the grid points
    moved by the float interpolation factors."""
    rel = flat_table(0.999, 1.001)
    na, nm, nr, ng = rel["n_a"], rel["n_mu0"], rel["n_r"], rel["n_g"]
    val = 0.5 * (alpha + beta * rel["a"][:, None] + gamma * rel["mu0"][None, :])
    t = np.repeat(val.reshape(na * nm), nr * ng).astype(np.float32)
    rel["trff1"] = t
    rel["trff2"] = t.copy()
    return rel


def affine_expected(rel, a, incl_deg, alpha=1.0, beta=0.3, gamma=0.5, swap_a=False):
    """alpha + beta a_f + gamma mu0_f: an affine field is reproduced exactly by
    bilinear interpolation, at the point a_f = a_i + f_a (a_i+1 - a_i) (same
    for mu0) with the factors f_a, f_mu0 computed in float as the reference
    does (src/Relprofile.cpp:171-177)."""
    ag, mg = rel["a"], rel["mu0"]
    af, mf = np.float32(a), np.float32(np.cos(np.radians(incl_deg)))
    ia = np.searchsorted(ag, af, side="right") - 1
    im = np.searchsorted(mg, mf, side="right") - 1
    fa = float(np.float32((af - ag[ia]) / (ag[ia + 1] - ag[ia])))
    fm = float(np.float32((mf - mg[im]) / (mg[im + 1] - mg[im])))
    if swap_a:  # a plausible mistake: the a weights exchanged
        fa = 1.0 - fa
    a_f = float(ag[ia]) + fa * (float(ag[ia + 1]) - float(ag[ia]))
    m_f = float(mg[im]) + fm * (float(mg[im + 1]) - float(mg[im]))
    return alpha + beta * a_f + gamma * m_f, (fa, fm)


def radial_table(u=0.01, v=1.0):
    """Flat window with trff_1 = trff_2 = (u + v r/1000)/2 at every table
    radius r: the node radii are the spin-lerped table radii and the fine-grid
    values lerp between two nodes (src/Relprofile.cpp:190-193, :249-300), so a
    correct mapping reproduces (u + v re/1000)/2 at every fine radius re."""
    rel = flat_table(0.999, 1.001)
    t = 0.5 * (u + v * rel["r"].astype(np.float64) / 1000.0)
    rel["trff1"] = np.repeat(t, rel["n_g"]).astype(np.float32)
    rel["trff2"] = rel["trff1"].copy()
    return rel


def radial_expected(rin, rout, u=0.01, v=1.0, q=3.0, nr=1000, node_r=None):
    """Flux ratio radial_table / flat table = emissivity-and-area weighted mean
    of (u + v re/1000) over the fine grid (re descending, weights re dr with
    half intervals at the ends, emissivity re^-q).  With node_r (descending
    node radii) the fine values are taken with the node bracket's lerp
    direction reversed: a plausible mistake the check must catch."""
    import oracle
    re = oracle.fine_radial_grid(rin, rout, nr)
    dr = np.empty_like(re)
    dr[0] = 0.5 * (re[0] - re[1])
    dr[-1] = 0.5 * (re[-2] - re[-1])
    dr[1:-1] = 0.5 * (re[:-2] - re[2:])
    w = re * dr * re ** -q
    x = re
    if node_r is not None:
        k = np.searchsorted(-node_r, -re, side="right") - 1  # node_r[k] >= re > node_r[k+1]
        x = node_r[k] + node_r[k + 1] - re
    return u + v / 1000.0 * np.sum(w * x) / np.sum(w)


def switch_lp_table(r_s=10.0):
    """The default lamp-post grids with fields that are exact under one of the
    two radius factors of rebin_emisprofile_on_radial_grid: below r_s the
    emission angle exceeds 75 deg and intensity is a power law, emission and
    incidence angles affine in log r (exact under the log factor); from r_s on
    the angle is below 75 deg and intensity exp(-lambda r), the angles affine
    in r (exact under the linear factor).  Same values at every spin and
    height."""
    rel, lp, T = tables()
    out = dict(lp)
    rad = lp["rad"].reshape(lp["n_a"], 1, lp["n_r"]).astype(np.float64)
    rad = np.broadcast_to(rad, (lp["n_a"], lp["n_h"], lp["n_r"]))
    lo = rad < r_s
    out["intens"] = np.where(lo, rad ** -3.0, 5e-3 * np.exp(-5e-3 * rad)).astype(np.float32).ravel()
    out["del_"] = np.where(lo, 1.40 + 0.05 * np.log(rad), 0.3 + 9e-4 * rad).astype(np.float32).ravel()
    out["del_inc"] = np.where(lo, 0.2 + 0.03 * np.log(rad), 0.1 + 5e-4 * rad).astype(np.float32).ravel()
    return rel, out


def switch_fields(r, lo):
    return (np.where(lo, r ** -3.0, 5e-3 * np.exp(-5e-3 * r)),
            np.where(lo, 1.40 + 0.05 * np.log(r), 0.3 + 9e-4 * r),
            np.where(lo, 0.2 + 0.03 * np.log(r), 0.1 + 5e-4 * r))



def isco_snap_table(above=5e-4):
    """Flat window, innermost table radius ISCO (1 + above) at every spin node
    (inside the 1e-3 snap band of src/Relprofile.cpp:195-200) and a transfer
    function equal to the table radius: trff_1 = trff_2 = r/2."""
    rel = flat_table(0.999, 1.001)
    na, nm, nr, ng = rel["n_a"], rel["n_mu0"], rel["n_r"], rel["n_g"]
    r = rel["r"].reshape(na, nm, nr).astype(np.float64).copy()
    for ia in range(na):
        r[ia, :, -1] = oracle.lib.orc_kerr_rms(float(rel["a"][ia])) * (1 + above)
    rel["r"] = r.astype(np.float32).ravel()
    t = 0.5 * rel["r"].astype(np.float64)
    rel["trff1"] = np.repeat(t, ng).astype(np.float32)
    rel["trff2"] = rel["trff1"].copy()
    return rel


def write_lptb(path, lp):
    """Writes a lamp-post table dict (tables_view layout) in the library's RXTB
    format: int32 magic 'RXTB', kind 2, n_a, n_h, n_r, then the float32 arrays
    a, h, rad, intens, del, del_inc."""
    hdr = np.array([0x42545852, 2, lp["n_a"], lp["n_h"], lp["n_r"]], dtype=np.int32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        for k in ("a", "h", "rad", "intens", "del_", "del_inc"):
            f.write(np.ascontiguousarray(lp[k], dtype=np.float32).tobytes())



# ---- FITS writer (binary tables, big-endian float32 columns) ----------------
# An independent writer of the layout the reference's loaders read
# (src/reltable.c:121-304 relline, :313-448 lamp post), so that the library's
# FITS loaders are checked against a file they did not write.

def _fits_card(key, value=None):
    if value is None:
        s = f"{key:<8}"
    elif isinstance(value, bool):
        s = f"{key:<8}= {'T' if value else 'F':>20}"
    elif isinstance(value, int):
        s = f"{key:<8}= {value:>20}"
    else:
        s = f"{key:<8}= '{str(value):<8}'"
    return s.ljust(80)[:80]


def _fits_header(cards):
    h = "".join(cards) + "END".ljust(80)
    return (h + " " * ((-len(h)) % 2880)).encode("ascii")


def _fits_bintable(name, cols):
    """cols: list of (column name, array of shape (rows, repeat)[, tform]) with
    tform "E" (float32, the default), "D" (float64), "J" (int32) or "<w>A"
    (strings of width w)."""
    prepared = []
    for c in cols:
        cname, v = c[0], c[1]
        form = c[2] if len(c) > 2 else "E"
        if form.endswith("A"):
            w = int(form[:-1])
            data = np.array([s.ljust(w)[:w].encode("ascii") for s in v], dtype=f"S{w}").reshape(len(v), 1)
            prepared.append((cname, form, data, w))
        else:
            dt = {"E": ">f4", "J": ">i4", "D": ">f8"}[form]
            a = np.asarray(v, dtype=dt).reshape(len(v), -1)
            prepared.append((cname, f"{a.shape[1]}{form}", a, a.itemsize * a.shape[1]))
    nrows = prepared[0][2].shape[0]
    assert all(d.shape[0] == nrows for _, _, d, _ in prepared)
    rowbytes = sum(w for _, _, _, w in prepared)
    cards = [_fits_card("XTENSION", "BINTABLE"), _fits_card("BITPIX", 8), _fits_card("NAXIS", 2),
             _fits_card("NAXIS1", rowbytes), _fits_card("NAXIS2", nrows), _fits_card("PCOUNT", 0),
             _fits_card("GCOUNT", 1), _fits_card("TFIELDS", len(prepared))]
    for i, (cname, form, _, _) in enumerate(prepared):
        cards += [_fits_card(f"TTYPE{i + 1}", cname), _fits_card(f"TFORM{i + 1}", form)]
    cards.append(_fits_card("EXTNAME", name))
    rows = bytearray()
    for r in range(nrows):
        for _, _, d, _ in prepared:
            rows += d[r].tobytes()
    data = bytes(rows)
    return _fits_header(cards) + data + b"\0" * ((-len(data)) % 2880)


def _fits_primary():
    return _fits_header([_fits_card("SIMPLE", True), _fits_card("BITPIX", 8), _fits_card("NAXIS", 0),
                         _fits_card("EXTEND", True)])


def write_fits_reltable(path, rel):
    """The relline table (tables_view layout) as rel_table_v0.5a.fits lays it
    out: extensions "a" and "mu0", then node (i, j) as HDU i*n_mu0 + j + 4 named
    "<i+1>_<j+1>" with rows r, gmin, gmax and vectors trff1, trff2, cosne1, cosne2."""
    na, nm, nr, ng = rel["n_a"], rel["n_mu0"], rel["n_r"], rel["n_g"]
    out = [_fits_primary(), _fits_bintable("a", [("a", rel["a"])]), _fits_bintable("mu0", [("mu0", rel["mu0"])])]
    r = rel["r"].reshape(na, nm, nr)
    gl, gh = rel["gmin"].reshape(na, nm, nr), rel["gmax"].reshape(na, nm, nr)
    t1, t2 = rel["trff1"].reshape(na, nm, nr, ng), rel["trff2"].reshape(na, nm, nr, ng)
    c1, c2 = rel["cosne1"].reshape(na, nm, nr, ng), rel["cosne2"].reshape(na, nm, nr, ng)
    for i in range(na):
        for j in range(nm):
            out.append(_fits_bintable(f"{i + 1}_{j + 1}", [("r", r[i, j]), ("gmin", gl[i, j]), ("gmax", gh[i, j]),
                                                            ("trff1", t1[i, j]), ("trff2", t2[i, j]),
                                                            ("cosne1", c1[i, j]), ("cosne2", c2[i, j])]))
    with open(path, "wb") as f:
        for blk in out:
            f.write(blk)


def write_fits_lptable(path, lp):
    """The lamp-post table as rel_lp_table_v0.5b.fits lays it out: extension
    "I_h", one row per spin with a, hgrid, r, and h<k>, del<k>, del_inc<k> for
    every height k."""
    na, nh, nr = lp["n_a"], lp["n_h"], lp["n_r"]
    inten = lp["intens"].reshape(na, nh, nr)
    de, di = lp["del_"].reshape(na, nh, nr), lp["del_inc"].reshape(na, nh, nr)
    cols = [("a", lp["a"]), ("hgrid", lp["h"].reshape(na, nh)), ("r", lp["rad"].reshape(na, nr))]
    for k in range(nh):
        cols += [(f"h{k + 1}", inten[:, k]), (f"del{k + 1}", de[:, k]), (f"del_inc{k + 1}", di[:, k])]
    with open(path, "wb") as f:
        f.write(_fits_primary())
        f.write(_fits_bintable("I_h", cols))


def draw_wide_params(model, rng):
    """Parameter arrays over the full lmodel.dat ranges, including values the
    reference rejects (Rout <= Rin) or rescales (negative Rin, Rout, h)."""
    a = rng.uniform(-0.998, 0.998)
    incl = rng.uniform(3.0, 87.0)
    rin = rng.choice([-1.0, -rng.uniform(1.0, 5.0), rng.uniform(1.0, 60.0)])
    rout = rng.choice([400.0, -rng.uniform(10.0, 300.0), rng.uniform(10.0, 1000.0)])
    limb = float(rng.integers(0, 3))
    if model == rx.MODEL_RELLINE:
        return [rng.uniform(1, 10), rng.uniform(-2, 6), rng.uniform(-2, 6), rng.uniform(1, 100), a, incl,
                rin, rout, rng.uniform(0, 1), limb]
    if model in (rx.MODEL_RELCONV, rx.MODEL_RELXILL_CONV):
        return [rng.uniform(-2, 6), rng.uniform(-2, 6), rng.uniform(1, 100), a, incl, rin, rout, limb]
    if model == rx.MODEL_RELLINELP:
        return [rng.uniform(1, 10), rng.uniform(-10, 120), a, incl, rin, rout, rng.uniform(0, 1), limb,
                rng.uniform(1, 3.4), float(rng.integers(0, 2))]
    return [rng.uniform(-10, 120), rng.uniform(-0.5, 1.0), a, incl, rin, rout, limb, rng.uniform(1, 3.4),
            float(rng.integers(0, 2))]


XILL_NAMES = {0: "Gamma", 1: "A_Fe", 2: "logXi", 3: "Ecut", 4: "Dens", 7: "Incl"}


def write_fits_xilltable(path, x):
    """A xillver table (xilltable_view layout) as an OGIP table model file like
    xillver-a-Ec5.fits: PARAMETERS (NAME, METHOD, INITIAL, DELTA, MINIMUM,
    BOTTOM, TOP, MAXIMUM, NUMBVALS, VALUE), ENERGIES (ENERG_LO, ENERG_HI),
    SPECTRA (PARAMVAL, INTPSPEC), the reader's column numbers 1, 9, 10 and 2."""
    n_par, n_val = x["n_par"], list(x["n_val"])
    nmax = max(n_val)
    vals = np.zeros((n_par, nmax), dtype=np.float32)
    for k in range(n_par):
        vals[k, :n_val[k]] = x["val"][k]
    first = np.array([x["val"][k][0] for k in range(n_par)], dtype=np.float32)
    last = np.array([x["val"][k][-1] for k in range(n_par)], dtype=np.float32)
    params = _fits_bintable("PARAMETERS", [
        ("NAME", [XILL_NAMES[p] for p in x["pid"]], "12A"), ("METHOD", np.zeros(n_par), "J"),
        ("INITIAL", first), ("DELTA", np.full(n_par, 0.01)), ("MINIMUM", first), ("BOTTOM", first),
        ("TOP", last), ("MAXIMUM", last), ("NUMBVALS", n_val, "J"), ("VALUE", vals)])
    energies = _fits_bintable("ENERGIES", [("ENERG_LO", x["elo"]), ("ENERG_HI", x["ehi"])])
    spec = np.asarray(x["spec"], dtype=np.float32).reshape(-1, x["n_ener"])
    grid = np.stack(np.meshgrid(*[x["val"][k] for k in range(n_par)], indexing="ij"), -1).reshape(-1, n_par)
    spectra = _fits_bintable("SPECTRA", [("PARAMVAL", grid[:spec.shape[0]]), ("INTPSPEC", spec)])
    with open(path, "wb") as f:
        for blk in (_fits_primary(), params, energies, spectra):
            f.write(blk)


def xill_axes(n_gam=4, n_afe=3, n_lxi=5, n_ect=4, n_incl=6):
    """Parameter axes of a small xillver table (Gamma, A_Fe, logXi, Ecut, Incl)."""
    return [np.linspace(1.0, 3.4, n_gam), np.geomspace(0.5, 10.0, n_afe), np.linspace(0.0, 4.7, n_lxi),
            np.geomspace(5.0, 1000.0, n_ect),
            np.degrees(np.arccos(1.0 - (np.arange(n_incl) + 0.5) / n_incl))]


AFFINE_XILL_C = (0.7, 0.21, 0.037, 0.083, 4.1e-4, 0.059)


def affine_xilltable(n_ener=300, c=AFFINE_XILL_C, **dims):
    """A xillver table whose stored spectra are 10^logXi times a shape g(E)
    times c0 + c1 Gamma + c2 A_Fe + c3 logXi + c4 Ecut + c5 k (k the index of
    the inclination): after the division by 10^logXi the values are affine in
    the four interpolated parameters, so the interpolation reproduces them."""
    ax = xill_axes(**dims)
    e = np.geomspace(0.07, 1000.0, n_ener + 1)
    g = 1e-3 * (0.5 * (e[1:] + e[:-1])) ** -1.0 * np.diff(e)
    G, A, L, C, I = np.meshgrid(*ax, indexing="ij")
    K = np.broadcast_to(np.arange(len(ax[4])), G.shape)
    lin = c[0] + c[1] * G + c[2] * A + c[3] * L + c[4] * C + c[5] * K
    spec = (10.0 ** L * lin)[..., None] * g
    return dict(n_par=5, pid=[0, 1, 2, 3, 7], n_val=[len(a) for a in ax],
                val=[np.asarray(a, dtype=np.float32) for a in ax], n_ener=n_ener,
                elo=e[:-1].astype(np.float32), ehi=e[1:].astype(np.float32),
                spec=spec.reshape(-1, n_ener).astype(np.float32)), g


def affine_xill_expected(x, g, gam, afe, lxi, ect, c=AFFINE_XILL_C):
    """[n_incl, n_ener] spectra the interpolation must give (parameters clamped
    into the table)."""
    v = x["val"]
    clamp = lambda q, k: min(max(q, float(v[k][0])), float(v[k][-1]))
    gam, afe, lxi, ect = clamp(gam, 0), clamp(afe, 1), clamp(lxi, 2), clamp(ect, 3)
    k = np.arange(x["n_val"][4])
    lin = c[0] + c[1] * gam + c[2] * afe + c[3] * lxi + c[4] * ect + c[5] * k
    return lin[:, None] * g[None, :]


def affine_xilltable6(n_ener=300, c=AFFINE_XILL_C, c_dens=0.053, n_dens=4, **dims):
    """A six-parameter table, Gamma, A_Fe, logXi, Ecut, Dens (15-19), Incl:
    stored spectra 10^logXi 10^(logN - 15) g(E) times the affine form of
    affine_xilltable plus c_dens logN, so that after the loader's divisions
    the values are affine in all six parameters."""
    a5 = xill_axes(**dims)
    ax = a5[:4] + [np.linspace(15.0, 19.0, n_dens)] + a5[4:]
    e = np.geomspace(0.07, 1000.0, n_ener + 1)
    g = 1e-3 * (0.5 * (e[1:] + e[:-1])) ** -1.0 * np.diff(e)
    G, A, L, C, D, I = np.meshgrid(*ax, indexing="ij")
    K = np.broadcast_to(np.arange(len(ax[5])), G.shape)
    lin = c[0] + c[1] * G + c[2] * A + c[3] * L + c[4] * C + c[5] * K + c_dens * D
    spec = (10.0 ** L * 10.0 ** (D - 15.0) * lin)[..., None] * g
    return dict(n_par=6, pid=[0, 1, 2, 3, 4, 7], n_val=[len(a) for a in ax],
                val=[np.asarray(a, dtype=np.float32) for a in ax], n_ener=n_ener,
                elo=e[:-1].astype(np.float32), ehi=e[1:].astype(np.float32),
                spec=spec.reshape(-1, n_ener).astype(np.float32)), g


def xill_slice(x6, i):
    """The five-parameter table of the first-axis node i of a six-parameter
    table view (its rows as stored)."""
    n_val = list(x6["n_val"][1:])
    rows = int(np.prod(n_val))
    return dict(n_par=5, pid=list(x6["pid"][1:]), n_val=n_val, val=list(x6["val"][1:]), n_ener=x6["n_ener"],
                elo=x6["elo"], ehi=x6["ehi"], spec=np.asarray(x6["spec"])[i * rows:(i + 1) * rows])


def draw_relxill_params(rng):
    """relxill parameters over the lmodel.dat ranges (refl_frac of both signs)."""
    a = rng.uniform(-0.998, 0.998)
    return [rng.uniform(-1, 6), rng.uniform(-1, 6), rng.uniform(2, 100), a, rng.uniform(5, 85),
            rng.choice([-1.0, rng.uniform(1.0, 30.0)]), rng.uniform(50, 1000), rng.choice([0.0, rng.uniform(0, 2)]),
            rng.uniform(1.0, 3.4), rng.uniform(0.0, 4.7), rng.uniform(0.5, 10.0), rng.uniform(5.0, 1000.0),
            rng.choice([-1.0, rng.uniform(-5, 10)])]


def rrad_table(n_spin=3, n_rad=12, n_g=6, spins=(-0.5, 0.5, 0.998), c=0.05, shift=0, g0=1.0, spin_scale=None):
    """A returning-radiation table: zones log-spaced from each spin's ISCO to
    1000 rg; radiation emitted in zone j returns to zone j + shift with the
    fraction c (times spin_scale[k] for spin k); all of it in one energy-shift
    bin at g = g0 (the middle of n_g bins, n_g even: g0 is the centre of bin
    n_g/2 when gmin, gmax are placed so)."""
    spins = np.asarray(spins[:n_spin], dtype=np.float64)
    nr2 = n_rad * n_rad
    rlo = np.zeros((n_spin, n_rad))
    rhi = np.zeros((n_spin, n_rad))
    tf = np.zeros((n_spin, n_rad, n_rad))
    gmin = np.zeros((n_spin, n_rad, n_rad))
    gmax = np.zeros((n_spin, n_rad, n_rad))
    fg = np.zeros((n_spin, n_rad, n_rad, n_g))
    for k, a in enumerate(spins):
        r0 = oracle.lib.orc_kerr_rms(float(a))
        edges = r0 * (1000.0 / r0) ** (np.arange(n_rad + 1) / n_rad)
        rlo[k], rhi[k] = edges[:-1], edges[1:]
        for j in range(n_rad):
            i = j + shift
            if 0 <= i < n_rad:
                tf[k, i, j] = c * (spin_scale[k] if spin_scale is not None else 1.0)
        # bin m = n_g // 2 has its centre at g0: (m + 0.5)/n_g (gmax - gmin) + gmin = g0
        m = n_g // 2
        width = 0.4
        gmin[k] = g0 - (m + 0.5) / n_g * width
        gmax[k] = gmin[k] + width
        fg[k, :, :, m] = 1.0
    return dict(n_spin=n_spin, n_rad=n_rad, n_g=n_g, spin=spins, rlo=rlo.ravel(), rhi=rhi.ravel(),
                tf_r=tf.ravel(), gmin=gmin.ravel(), gmax=gmax.ravel(), frac_g=fg.ravel())


def write_fits_rradtable(path, t):
    """The returning-radiation table as table_returnRad_v20220301.fits lays it
    out: extension SPIN (column a), then FRAC01, ... with rlo, rhi, frac_e, tf_r,
    gmin, gmax, frac_g, f_ret (double columns; row = incident zone)."""
    ns, nr, ng = t["n_spin"], t["n_rad"], t["n_g"]
    out = [_fits_primary(), _fits_bintable("SPIN", [("a", np.asarray(t["spin"]), "D")])]
    rlo = np.asarray(t["rlo"]).reshape(ns, nr)
    rhi = np.asarray(t["rhi"]).reshape(ns, nr)
    tf = np.asarray(t["tf_r"]).reshape(ns, nr, nr)
    gl = np.asarray(t["gmin"]).reshape(ns, nr, nr)
    gh = np.asarray(t["gmax"]).reshape(ns, nr, nr)
    fg = np.asarray(t["frac_g"]).reshape(ns, nr, nr * ng)
    for k in range(ns):
        out.append(_fits_bintable(f"FRAC{k + 1:02d}", [
            ("rlo", rlo[k], "D"), ("rhi", rhi[k], "D"), ("frac_e", np.zeros((nr, nr)), "D"), ("tf_r", tf[k], "D"),
            ("gmin", gl[k], "D"), ("gmax", gh[k], "D"), ("frac_g", fg[k], "D"), ("f_ret", tf[k].sum(axis=1), "D")]))
    with open(path, "wb") as f:
        for blk in out:
            f.write(blk)


# ---- closed forms for the angle distribution, the red-wing midpoint rule and
# the lamp-post spin / height interpolation -----------------------------------

def dist_table(gmin=0.3, gmax=1.4, c1=0.7, c2=0.3):
    """Flat window [gmin, gmax] (the same at every node and radius), trff_1 =
    c1, trff_2 = c2, and emission-angle cosines that depend on the g* node
    only: branch 1 of node g has its cosine at the centre of angle bin g,
    branch 2 at the centre of bin n_g + g, of 2 n_g cos-uniform bins.
    Returns (table, n_incl)."""
    rel = flat_table(gmin, gmax)
    ng = rel["n_g"]
    nn = rel["n_a"] * rel["n_mu0"] * rel["n_r"]
    n_incl = 2 * ng
    mu1 = 1.0 - (np.arange(ng) + 0.5) / n_incl
    mu2 = 1.0 - (ng + np.arange(ng) + 0.5) / n_incl
    rel["trff1"] = np.full(nn * ng, c1, dtype=np.float32)
    rel["trff2"] = np.full(nn * ng, c2, dtype=np.float32)
    rel["cosne1"] = np.tile(mu1, nn).astype(np.float32)
    rel["cosne2"] = np.tile(mu2, nn).astype(np.float32)
    return rel, n_incl


def dist_expected(ng, gmin=0.3, gmax=1.4, c1=0.7, c2=0.3, shift=0, drop_dgstar=False, drop_g2=False):
    """Normalised angle distribution on dist_table().  The reference adds, per
    radius and (g* node j, branch k), r (2 pi g r)^2 trff_k d_gstar_j /
    sqrt(g*_j (1 - g*_j)) emis weight into bin floor(n (1 - mu)) (with g =
    g*_j (gmax - gmin) + gmin, src/Relprofile.cpp:908-938, :799-801).  With a
    radius-independent window, transfer function and cosines the radial sum
    r^3 emis weight is a common factor: bin j holds c1 W_j, bin n_g + j holds
    c2 W_j, W_j = d_gstar_j g_j^2 / sqrt(g*_j (1 - g*_j)), with the g* grid
    and its trapezoid widths of src/Relprofile.cpp:104-116 (half a step plus
    H at both ends); gmin, gmax, c1, c2 as the float32 table stores them.
    shift / drop_* give plausible mistakes."""
    gmin, gmax, c1, c2 = (float(np.float32(x)) for x in (gmin, gmax, c1, c2))
    H = 5e-3
    gs = H + (1 - 2 * H) / (ng - 1) * np.arange(ng)
    step = gs[1] - gs[0]
    d = np.full(ng, step)
    d[0] = d[-1] = 0.5 * step + H
    if drop_dgstar:
        d = np.ones(ng)
    g = gs * (gmax - gmin) + gmin
    W = d * (1.0 if drop_g2 else g ** 2) / np.sqrt(gs * (1 - gs))
    out = np.concatenate([c1 * W, c2 * W])
    out = np.roll(out, shift)
    return out / out.sum()


def midpoint_bins(e, gmin, gmax, c=1.0, rule="mid"):
    """Photons per bin of relline on flat_table(gmin, gmax, c) with the
    physical normalisation, for the bins whose g* range lies inside (H, 1-H)
    of a window below 0.95 (in line-energy units): there int_romb takes the
    midpoint rule f((lo+hi)/2) (hi - lo) (src/Relprofile.cpp:636-644), f =
    E^3 c / ((gmax - gmin) sqrt(g*(1-g*))) summed over both branches, and
    the normalised emissivity gives sum_r emis pi r dr / 2 = 1/4.  rule
    "quad" integrates f instead (what Romberg everywhere would give), "lo"
    takes the lower end: the mistakes the check must catch.  Returns
    (indices, values).  gmin, gmax as the float32 table stores them."""
    from scipy.integrate import quad
    gmin, gmax = float(np.float32(gmin)), float(np.float32(gmax))
    H = 5e-3
    f = lambda E: E ** 3 * c / ((gmax - gmin) * np.sqrt(((E - gmin) / (gmax - gmin)) * (1 - (E - gmin) / (gmax - gmin))))
    idx, val = [], []
    for j in range(len(e) - 1):
        lo, hi = e[j], e[j + 1]
        glo, ghi = (lo - gmin) / (gmax - gmin), (hi - gmin) / (gmax - gmin)
        if not (glo > H and ghi < 1 - H):
            continue
        if rule == "mid":
            v = f(0.5 * (lo + hi)) * (hi - lo)
        elif rule == "lo":
            v = f(lo) * (hi - lo)
        else:
            v = quad(f, lo, hi, epsabs=0, epsrel=1e-13)[0]
        idx.append(j)
        val.append(0.25 * v / (0.5 * (lo + hi)))
    return np.array(idx), np.array(val)


LP_INT = (1.0, 0.4, 0.01)     # intensity factor A = A0 + A1 a + A2 h
LP_DEL = (1.38, 0.03, 4e-4)   # emission angle D + s ln r (above 75 deg everywhere)
LP_DEL_S = 0.04
LP_INC = (0.2, 0.05, 1e-3)    # incidence angle Di + si ln r
LP_INC_S = 0.02


def affine_lp_table(unit=False):
    """A lamp-post table whose fields are affine in spin and height at every
    table radius: intensity A(a, h) r^-3, emission angle D(a, h) + s ln r,
    incidence angle Di(a, h) + si ln r.  The emission angle exceeds 75 deg
    everywhere, so rebin_emisprofile_on_radial_grid uses the log radius factor
    (src/Rellp.cpp:168-173), under which these r-dependences are exact; the
    radii are one log grid from 1 to 1000 rg for every spin.  Every spin has
    its own height grid (1.5 .. 500 rg with a spin-dependent spacing), so a
    height factor taken from the wrong spin is visible.  unit=True sets A = 1
    (the reference the intensity ratio is taken against)."""
    rel, lp, T = tables()
    na, nh, nr = lp["n_a"], lp["n_h"], lp["n_r"]
    out = dict(lp)
    a = lp["a"].astype(np.float64)
    t = np.arange(nh) / (nh - 1)
    hg = np.empty((na, nh))
    for ia in range(na):
        p = 1.0 + 0.35 * (ia % 4)
        hg[ia] = 1.5 * (500.0 / 1.5) ** (t ** p)
    hg = hg.astype(np.float32)
    rad = np.exp(np.log(1000.0) * np.arange(nr) / (nr - 1))
    rad[-1] = 1000.0
    rad = rad.astype(np.float32)
    A = np.asarray(a)[:, None, None] * 0 + 1.0 if unit else \
        LP_INT[0] + LP_INT[1] * a[:, None, None] + LP_INT[2] * hg[:, :, None].astype(np.float64)
    r = rad.astype(np.float64)[None, None, :]
    D = LP_DEL[0] + LP_DEL[1] * a[:, None, None] + LP_DEL[2] * hg[:, :, None].astype(np.float64)
    Di = LP_INC[0] + LP_INC[1] * a[:, None, None] + LP_INC[2] * hg[:, :, None].astype(np.float64)
    out["h"] = hg.ravel()
    out["rad"] = np.tile(rad, na)
    out["intens"] = np.broadcast_to(A * r ** -3.0, (na, nh, nr)).astype(np.float32).ravel()
    out["del_"] = (D + LP_DEL_S * np.log(r)).astype(np.float32).ravel()
    out["del_inc"] = (Di + LP_INC_S * np.log(r)).astype(np.float32).ravel()
    return out


def lp_factors(lp, a, h, mistake=None):
    """Spin and height brackets with the interpolation factors get_ipol_factor
    forms in float (value, grid and quotient in float, src/relutility.c:419-423;
    binary_search_float, :135-151), the height factor per bracketing spin on
    that spin's own grid (src/Rellp.cpp:225-241).  mistake "one_h" uses the
    lower spin's height bracket and factor for both spins, "swap_a" exchanges
    the spin weights."""
    ag = lp["a"].astype(np.float32)
    hg = lp["h"].reshape(lp["n_a"], lp["n_h"]).astype(np.float32)

    def bracket(grid, v):
        i = int(np.searchsorted(grid, v, side="right")) - 1
        i = min(max(i, 0), len(grid) - 2)
        f = np.float32(np.float32(v - grid[i]) / np.float32(grid[i + 1] - grid[i]))
        return i, float(f)
    ia, fa = bracket(ag, np.float32(a))
    hs = [bracket(hg[ia + k], np.float32(h)) for k in (0, 1)]
    if mistake == "one_h":
        hs[1] = hs[0]
    if mistake == "swap_a":
        fa = 1.0 - fa
    # the interpolated points: a moved by fa, h by each spin's factor
    a_f = float(ag[ia]) + fa * (float(ag[ia + 1]) - float(ag[ia]))
    h_f = [float(hg[ia + k][hs[k][0]]) + hs[k][1] * (float(hg[ia + k][hs[k][0] + 1]) - float(hg[ia + k][hs[k][0]]))
           for k in (0, 1)]
    return fa, a_f, h_f


def lp_affine_value(coef, lp, a, h, mistake=None):
    """c0 + c1 a + c2 h interpolated as interpol_lptable does
    ((1 - f_a) lerp_h(spin lo) + f_a lerp_h(spin hi), src/Rellp.cpp:219-247):
    bilinear interpolation reproduces an affine field, at the points moved by
    the float factors."""
    fa, a_f, h_f = lp_factors(lp, a, h, mistake)
    return coef[0] + coef[1] * a_f + coef[2] * ((1 - fa) * h_f[0] + fa * h_f[1])


def aberrate(delta, beta):
    """Emission angle (from the direction to the black hole) in the rest frame
    of a source moving away from the disc with speed beta, from the static
    angle: cos d' = (cos d + beta) / (1 + beta cos d) (relativistic
    aberration)."""
    c = np.cos(delta)
    return np.arccos((c + beta) / (1 + beta * c))


def refl_frac_expected(lp, a, h, beta, rin, rout, nr=1000):
    """Reflection fractions of calc_refl_frac (src/Rellp.cpp:36-89) on
    affine_lp_table(): the emission angles at the fine radius the reference's
    bisections land on for Rin (index nr - 2) and Rout (index 0) of the
    descending fine grid, D + s ln r there; the largest table angle D + s
    ln 1000 (not below pi/2); for beta > 1e-6 all three taken to the source's
    rest frame.  f_bh = (1 - cos d_bh)/2, f_ad = (cos d_bh - cos d_ad)/2,
    f_inf = (1 + cos d_max)/2, refl_frac = f_ad / f_inf.  Returns [refl_frac,
    f_bh, f_ad, f_inf, f_inf_rest]."""
    import oracle
    D = lp_affine_value(LP_DEL, lp, a, h)
    re = oracle.fine_radial_grid(rin, rout, nr)
    d_bh = D + LP_DEL_S * np.log(re[nr - 2])
    d_ad = D + LP_DEL_S * np.log(re[0])
    d_max = max(D + LP_DEL_S * np.log(1000.0), np.pi / 2)
    f_inf_rest = 0.5 * (1 + np.cos(d_max))
    if beta > 1e-6:
        d_bh, d_ad, d_max = aberrate(d_bh, beta), aberrate(d_ad, beta), aberrate(d_max, beta)
    f_bh = 0.5 * (1 - np.cos(d_bh))
    f_ad = 0.5 * (np.cos(d_bh) - np.cos(d_ad))
    f_inf = 0.5 * (1 + np.cos(d_max))
    return np.array([f_ad / f_inf, f_bh, f_ad, f_inf, f_inf_rest])
