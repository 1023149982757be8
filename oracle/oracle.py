"""ctypes bindings of the FP64 CPU oracle (oracle/psoracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the parity checker / CPU baseline.
The product library (``powspec_amd``) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_math.so")

NGP, CIC, TSC, PCS = 0, 1, 2, 3
MAX_ELL = 6

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


class Conf(C.Structure):
    """pso_conf: the CONF fields the path reads (src/load_conf.h:40-93)."""

    _fields_ = [
        ("ndata", C.c_int), ("issim", C.c_int), ("gsize", C.c_int),
        ("assign", C.c_int), ("intlace", C.c_int), ("npole", C.c_int),
        ("poles", C.c_int * (MAX_ELL + 1)), ("has_bsize", C.c_int),
        ("bsize", C.c_double * 3), ("bpad", C.c_double * 3),
        ("los", C.c_double * 3), ("logscale", C.c_int),
        ("kmin", C.c_double), ("kmax", C.c_double), ("kbin", C.c_double),
        ("isauto", C.c_int * 2), ("iscross", C.c_int),
    ]


class Cata(C.Structure):
    _fields_ = [
        ("num", C.c_int),
        ("data", C.c_void_p * 2), ("ndata", C.c_size_t * 2),
        ("rand", C.c_void_p * 2), ("nrand", C.c_size_t * 2),
        ("wdata", C.c_double * 2), ("wrand", C.c_double * 2),
        ("alpha", C.c_double * 2), ("shot", C.c_double * 2),
        ("norm", C.c_double * 2),
    ]


class Result(C.Structure):
    _fields_ = [
        ("cap", C.c_int), ("nbin", C.c_int), ("nl", C.c_int),
        ("k", C.c_void_p), ("kedge", C.c_void_p), ("km", C.c_void_p),
        ("cnt", C.c_void_p), ("lcnt", C.c_void_p), ("pl", C.c_void_p * 2),
        ("xpl", C.c_void_p), ("bsize", C.c_double * 3),
        ("bmin", C.c_double * 3), ("cmax", C.c_double * 3),
    ]


_lib = None
_ref = None


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and _ref/libref_math.so when the reference is
    present) with oracle/Makefile."""
    out = subprocess.run(["make", "-C", HERE, "all", "ref"], capture_output=True,
                         text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.pso_assign.argtypes = [C.c_int, _dp, C.c_size_t, C.c_int, _dp, _dp, _dp]
        L.pso_assign.restype = None
        L.pso_coord_bound.argtypes = [_dp, C.c_size_t, _dp, _dp]
        L.pso_def_box.argtypes = [_dp, _dp, C.c_int, C.c_void_p, _dp, _dp, _dp]
        L.pso_def_box.restype = C.c_int
        L.pso_shift_cat.argtypes = [_dp, C.c_size_t, _dp, _dp]
        L.pso_alias_corr.argtypes = [C.c_int, C.c_int, C.c_double]
        L.pso_alias_corr.restype = C.c_double
        L.pso_kbins.argtypes = [C.c_int, _dp, C.c_int, C.c_double, C.c_double,
                                C.c_double, C.c_int, C.c_void_p, C.c_void_p]
        L.pso_kbins.restype = C.c_int
        L.pso_precomp.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                  _dp, C.c_double, _dp, _u64p, _dp]
        L.pso_fft_r2c.argtypes = [C.c_int, _dp, _dp]
        L.pso_fft_c2r.argtypes = [C.c_int, _dp, _dp]
        L.pso_interlace.argtypes = [C.c_int, _dp, _dp]
        L.pso_count_sim.argtypes = [C.c_int, _dp, C.c_int, C.c_int, _dp, C.c_double,
                                    _dp, C.c_int, C.POINTER(C.c_int), _dp, _dp, _dp,
                                    _dp, _dp]
        L.pso_count_survey.argtypes = [C.c_int, _dp, C.c_int, C.c_int, _dp, C.c_double,
                                       _dp, _dp, _dp, _dp]
        L.pso_ylm.argtypes = [C.c_int, C.c_int] + [C.c_double] * 4
        L.pso_ylm.restype = C.c_double
        L.pso_legpoly.argtypes = [C.c_int, C.c_double]
        L.pso_legpoly.restype = C.c_double
        L.pso_dens_kl.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.pso_run.argtypes = [C.POINTER(Conf), C.POINTER(Cata), C.POINTER(Result)]
        L.pso_run.restype = C.c_int
        L.pso_last_timings.argtypes = [_dp]
        L.pso_num_threads.restype = C.c_int
        L.pso_legauss.argtypes = [C.c_int, _dp, _dp, _dp]
        L.pso_cmvdist.argtypes = [C.c_double] * 4 + [C.c_int, C.c_double]
        L.pso_cmvdist.restype = C.c_double
        L.pso_cmvdist_order.argtypes = [C.c_double] * 7
        L.pso_cmvdist_order.restype = C.c_int
        L.pso_spline_ypp.argtypes = [_dp, _dp, C.c_size_t, _dp]
        L.pso_spline_ypp.restype = C.c_int
        L.pso_spline_dist.argtypes = [_dp, _dp, _dp, C.c_size_t, C.c_double]
        L.pso_spline_dist.restype = C.c_double
        L.pso_cnvt.argtypes = [_dp, C.c_size_t] + [C.c_double] * 4 + [C.c_int, C.c_void_p,
                                                                     C.c_void_p, C.c_void_p,
                                                                     C.c_size_t]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own math headers compiled as-is, or None when absent."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        R.ref_ylm.argtypes = [C.c_int, C.c_int] + [C.c_double] * 4
        R.ref_ylm.restype = C.c_double
        R.ref_legpoly.argtypes = [C.c_int, C.c_double]
        R.ref_legpoly.restype = C.c_double
        _ref = R
    return _ref


def _v3(x):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(x, np.float64), (3,)))


# ---- building blocks -------------------------------------------------------

def assign(scheme, data, Ng, bmin, L, rho=None):
    data = np.ascontiguousarray(data, np.float64)
    if rho is None:
        rho = np.zeros((Ng, Ng, Ng))
    lib().pso_assign(scheme, data, len(data), Ng, _v3(bmin), _v3(L), rho)
    return rho


def coord_bound(data, mn=None, mx=None):
    data = np.ascontiguousarray(data, np.float64)
    mn = np.full(3, np.finfo(float).max) if mn is None else mn
    mx = np.full(3, -np.finfo(float).max) if mx is None else mx
    lib().pso_coord_bound(data, len(data), mn, mx)
    return mn, mx


def def_box(mn, mx, issim, bsize, bpad):
    size = np.zeros(3)
    bmin = np.zeros(3)
    bs = None if bsize is None else _v3(bsize)
    err = lib().pso_def_box(_v3(mn), _v3(mx), int(issim),
                            None if bs is None else bs.ctypes.data, _v3(bpad),
                            size, bmin)
    return err, size, bmin


def shift_cat(data, smin, L):
    lib().pso_shift_cat(data, len(data), _v3(smin), _v3(L))


def alias_corr(intlace, scheme, pik):
    return lib().pso_alias_corr(int(intlace), scheme, float(pik))


def kbins(Ng, L, logscale, kmin, kmax, dk, cap=1 << 16):
    kedge = np.zeros(cap + 1)
    kcen = np.zeros(cap)
    nb = lib().pso_kbins(Ng, _v3(L), int(logscale), kmin, kmax, dk, cap,
                         kedge.ctypes.data, kcen.ctypes.data)
    if nb < 0:
        return nb, None, None
    return nb, kedge[: nb + 1].copy(), kcen[:nb].copy()


def precomp(Ng, L, intlace, scheme, logscale, kedge, dk):
    nbin = len(kedge) - 1
    alias = np.zeros((Ng, Ng, Ng // 2 + 1))
    cnt = np.zeros(nbin, np.uint64)
    km = np.zeros(nbin)
    lib().pso_precomp(Ng, _v3(L), int(intlace), scheme, int(logscale), nbin,
                      np.ascontiguousarray(kedge), dk, alias, cnt, km)
    return alias, cnt, km


def fft_r2c(rho):
    Ng = rho.shape[0]
    out = np.zeros((Ng, Ng, Ng // 2 + 1), np.complex128)
    lib().pso_fft_r2c(Ng, np.ascontiguousarray(rho, np.float64), out.view(np.float64))
    return out


def fft_c2r(F, Ng):
    out = np.zeros((Ng, Ng, Ng))
    lib().pso_fft_c2r(Ng, np.ascontiguousarray(F, np.complex128).view(np.float64), out)
    return out


def interlace(F0, F1):
    F0 = np.ascontiguousarray(F0, np.complex128).copy()
    lib().pso_interlace(F0.shape[0], F0.view(np.float64),
                        np.ascontiguousarray(F1, np.complex128).view(np.float64))
    return F0


def count_sim(Ng, L, logscale, kedge, dk, los, poles, alias, F1, F2):
    nbin = len(kedge) - 1
    nl = len(poles)
    pl = np.zeros(nl * nbin)
    lcnt = np.zeros(nl * nbin)
    parr = (C.c_int * nl)(*poles)
    lib().pso_count_sim(Ng, _v3(L), int(logscale), nbin, np.ascontiguousarray(kedge),
                        dk, _v3(los), nl, parr, np.ascontiguousarray(alias),
                        np.ascontiguousarray(F1).view(np.float64),
                        np.ascontiguousarray(F2).view(np.float64), pl, lcnt)
    return pl.reshape(nl, nbin), lcnt.reshape(nl, nbin)


def count_survey(Ng, L, logscale, kedge, dk, alias, F1, F2):
    nbin = len(kedge) - 1
    pl = np.zeros(nbin)
    lib().pso_count_survey(Ng, _v3(L), int(logscale), nbin, np.ascontiguousarray(kedge),
                           dk, np.ascontiguousarray(alias),
                           np.ascontiguousarray(F1).view(np.float64),
                           np.ascontiguousarray(F2).view(np.float64), pl)
    return pl


def ylm(ell, m, ct, st, cp, sp):
    return lib().pso_ylm(ell, m, ct, st, cp, sp)


def legpoly(ell, x):
    return lib().pso_legpoly(ell, x)


def dens_kl(ell, Fr, L, smin_g, alias):
    Ng = Fr.shape[0]
    nc = (Ng, Ng, Ng // 2 + 1)
    Frl = np.zeros_like(Fr)
    Fka = np.zeros(nc, np.complex128)
    Fkl = np.zeros(nc, np.complex128)
    lib().pso_dens_kl(ell, Ng, _v3(L), _v3(smin_g), np.ascontiguousarray(Fr), alias,
                      Frl, Fka.view(np.float64), Fkl.view(np.float64))
    return Fkl


# ---- the whole path --------------------------------------------------------

def make_conf(**kw) -> Conf:
    """Build a Conf from keyword arguments (names as in pso_conf)."""
    c = Conf()
    poles = kw.pop("poles", [0, 2, 4])
    c.npole = len(poles)
    for i, p in enumerate(poles):
        c.poles[i] = p
    c.ndata = kw.pop("ndata", 1)
    c.issim = int(kw.pop("issim", True))
    c.gsize = kw.pop("gsize")
    c.assign = kw.pop("assign", CIC)
    c.intlace = int(kw.pop("intlace", False))
    bsize = kw.pop("bsize", None)
    c.has_bsize = int(bsize is not None)
    if bsize is not None:
        c.bsize[:] = list(_v3(bsize))
    c.bpad[:] = list(_v3(kw.pop("bpad", 0.02)))
    c.los[:] = list(_v3(kw.pop("los", [0, 0, 1])))
    c.logscale = int(kw.pop("logscale", False))
    c.kmin = kw.pop("kmin", 0.0)
    c.kmax = kw.pop("kmax", -1.0)
    c.kbin = kw.pop("kbin")
    ia = kw.pop("isauto", [True, False])
    c.isauto[0], c.isauto[1] = int(ia[0]), int(ia[1])
    c.iscross = int(kw.pop("iscross", False))
    if kw:
        raise TypeError(f"unknown conf keys {sorted(kw)}")
    return c


def run(conf: Conf, data, rand=None, wsums=None, cap=1 << 15):
    """genr_mesh + powspec on catalogue(s).

    data/rand: list (one per catalogue) of (n,4) float64 arrays (x,y,z,w); they
    are copied (the oracle shifts coordinates in place like the reference).
    wsums: for surveys, list of dicts with wdata, wrand, alpha, shot, norm.
    Returns a dict of numpy arrays.
    """
    num = len(data)
    keep = []
    ct = Cata()
    ct.num = num
    for c in range(num):
        d = np.ascontiguousarray(data[c], np.float64).copy()
        keep.append(d)
        ct.data[c] = d.ctypes.data
        ct.ndata[c] = len(d)
        ct.wdata[c] = float(d[:, 3].sum()) if len(d) else 0.0
        if rand is not None:
            r = np.ascontiguousarray(rand[c], np.float64).copy()
            keep.append(r)
            ct.rand[c] = r.ctypes.data
            ct.nrand[c] = len(r)
        if wsums is not None:
            s = wsums[c]
            ct.wdata[c] = s["wdata"]
            ct.wrand[c] = s.get("wrand", 0.0)
            ct.alpha[c] = s.get("alpha", 0.0)
            ct.shot[c] = s.get("shot", 0.0)
            ct.norm[c] = s.get("norm", 0.0)
    nl = conf.npole
    arrs = {
        "k": np.zeros(cap), "kedge": np.zeros(cap + 1), "km": np.zeros(cap),
        "cnt": np.zeros(cap, np.uint64), "lcnt": np.zeros(nl * cap),
        "pl0": np.zeros(nl * cap), "pl1": np.zeros(nl * cap), "xpl": np.zeros(nl * cap),
    }
    res = Result()
    res.cap = cap
    res.k = arrs["k"].ctypes.data
    res.kedge = arrs["kedge"].ctypes.data
    res.km = arrs["km"].ctypes.data
    res.cnt = arrs["cnt"].ctypes.data
    res.lcnt = arrs["lcnt"].ctypes.data
    res.pl[0] = arrs["pl0"].ctypes.data
    res.pl[1] = arrs["pl1"].ctypes.data
    res.xpl = arrs["xpl"].ctypes.data
    err = lib().pso_run(C.byref(conf), C.byref(ct), C.byref(res))
    if err:
        return {"err": err}
    nb = res.nbin
    out = {
        "err": 0, "nbin": nb, "k": arrs["k"][:nb].copy(),
        "kedge": arrs["kedge"][: nb + 1].copy(), "km": arrs["km"][:nb].copy(),
        "cnt": arrs["cnt"][:nb].copy(),
        "lcnt": arrs["lcnt"][: nl * nb].reshape(nl, nb).copy(),
        "pl": [arrs["pl0"][: nl * nb].reshape(nl, nb).copy(),
               arrs["pl1"][: nl * nb].reshape(nl, nb).copy()],
        "xpl": arrs["xpl"][: nl * nb].reshape(nl, nb).copy(),
        "bsize": np.array(res.bsize[:]), "bmin": np.array(res.bmin[:]),
        "shot": [ct.shot[0], ct.shot[1]], "norm": [ct.norm[0], ct.norm[1]],
        "wdata": [ct.wdata[0], ct.wdata[1]],
    }
    t = np.zeros(4)
    lib().pso_last_timings(t)
    out["timings"] = t
    out["data_shifted"] = keep
    return out


def survey_sums(wc_d, wfkp_d, nz_d, wc_r, wfkp_r, nz_r):
    """Per-object weights and catalogue sums of a survey pair, following
    io/read_ascii.c:875-903 (w = w_c * w_fkp; sums of w_c, w^2 and
    w_c * w_fkp^2 * n) and src/read_cata.c:165-183 (alpha, shot, norm).
    Returns (w_d, w_r, dict)."""
    wd = wc_d * wfkp_d
    wr = wc_r * wfkp_r
    sw = [float(np.sum(wc_d)), float(np.sum(wc_r))]
    s2 = [float(np.sum(wd * wd)), float(np.sum(wr * wr))]
    s2n = [float(np.sum(wc_d * wfkp_d * wfkp_d * nz_d)),
           float(np.sum(wc_r * wfkp_r * wfkp_r * nz_r))]
    if sw[0] == 0 or s2[0] == 0 or sw[1] == 0 or s2[1] == 0:
        raise ValueError("invalid completeness or FKP weights")
    alpha = sw[0] / sw[1]
    shot = s2[0] + alpha * alpha * s2[1]
    if s2n[0] == 0:
        norm = alpha * s2n[1]
    elif s2n[1] == 0:
        norm = s2n[0]
    else:
        norm = alpha * s2n[1]
    return wd, wr, {"wdata": sw[0], "wrand": sw[1], "alpha": alpha, "shot": shot,
                    "norm": norm}


# ---- survey coordinate conversion (src/cnvt_coord.c) -----------------------

def legauss(n):
    """(positive abscissas, their weights, weight of x = 0 or 0)."""
    x = np.zeros(17)
    w = np.zeros(17)
    w0 = np.zeros(1)
    lib().pso_legauss(n, x, w, w0)
    return x[: n // 2].copy(), w[: n // 2].copy(), float(w0[0])


def eos_widx(eos_w):
    """Power index of the dark-energy term (src/cnvt_coord.c:322)."""
    return 0.0 if eos_w == -1 else 3 * (1 + eos_w)


def cmvdist(om, ol, ok, widx, n, z):
    return lib().pso_cmvdist(om, ol, ok, widx, n, z)


def cmvdist_order(om, ol, ok, widx, err, zmin, zmax):
    return lib().pso_cmvdist_order(om, ol, ok, widx, err, zmin, zmax)


def spline_ypp(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    ypp = np.zeros_like(x)
    if lib().pso_spline_ypp(x, y, len(x), ypp):
        raise ValueError("spline needs two or more samples")
    return ypp


def spline_dist(zs, ds, ypp, z):
    zs, ds, ypp = (np.ascontiguousarray(a, np.float64) for a in (zs, ds, ypp))
    return lib().pso_spline_dist(zs, ds, ypp, len(zs), z)


def cnvt(data, om=0.31, ol=0.69, ok=0.0, eos_w=-1.0, order=None, err=1e-8,
         zs=None, ds=None):
    """(RA, Dec, z, w) records -> (x, y, z, w), a converted copy."""
    d = np.ascontiguousarray(data, np.float64).copy()
    widx = eos_widx(eos_w)
    if zs is None:
        if order is None:
            order = cmvdist_order(om, ol, ok, widx, err, d[:, 2].min(), d[:, 2].max())
        lib().pso_cnvt(d, len(d), om, ol, ok, widx, order, None, None, None, 0)
    else:
        zs = np.ascontiguousarray(zs, np.float64)
        ds = np.ascontiguousarray(ds, np.float64)
        ypp = spline_ypp(zs, ds)
        lib().pso_cnvt(d, len(d), om, ol, ok, widx, 0, zs.ctypes.data, ds.ctypes.data,
                       ypp.ctypes.data, len(zs))
    return d
