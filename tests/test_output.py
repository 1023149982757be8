"""Output files (psh_write_pk) against the reference's save_res format
(src/save_res.c:35-230: '%.10lg' numbers, '%lg' box, '%zu' counts, '#'
comments), rendered independently here with Python's C-style formatting.
CPU only: the writer is host code."""
import ctypes as C
import os

import numpy as np
import pytest

import powspec_amd as P

NAMES = ["NGP", "CIC", "TSC", "PCS"]


def _render_cols(nbin, poles, k, kedge, km, cnt, pl):
    s = "# kcen(1) kmin(2) kmax(3) kavg(4) nmod(5)"
    s += "".join(" P_%d(%d)" % (p, l + 6) for l, p in enumerate(poles)) + "\n"
    for i in range(nbin):
        s += "%.10g %.10g %.10g %.10g %d" % (k[i], kedge[i], kedge[i + 1], km[i], cnt[i])
        s += "".join(" %.10g" % pl[l][i] for l in range(len(poles))) + "\n"
    return s


def _render_mesh(meta):
    b, m = meta["bsize"], meta["bmin"]
    s = "# Box size: [%g, %g, %g]\n" % tuple(b)
    s += "# Box boundaries: [[%g,%g], [%g,%g], [%g,%g]]\n" % (
        m[0], m[0] + b[0], m[1], m[1] + b[1], m[2], m[2] + b[2])
    s += "# Grid size: %d , assignment scheme: %s , grid interlacing: %s\n" % (
        meta["Ng"], NAMES[meta["assign"]], "enabled" if meta["intlace"] else "disabled")
    return s


def _case(issim, seed):
    rng = np.random.default_rng(seed)
    nbin, poles = 7, [0, 2, 4]
    kedge = np.cumsum(np.r_[0.0, rng.uniform(0.001, 0.02, nbin)])
    vals = np.array([1e-5, -123456.789012345, 0.0, 3.0, 2.5e12, -7.25e-9, 1.0 / 3])
    pk = dict(nbin=nbin, poles=poles, k=0.5 * (kedge[1:] + kedge[:-1]), kedge=kedge,
              km=kedge[:-1] + 1e-4, cnt=rng.integers(0, 10 ** 9, nbin).astype(np.uint64),
              pl=[np.stack([vals * (1 + l + 10 * c) for l in range(3)]) for c in range(2)],
              xpl=np.stack([vals * (0.1 - l) for l in range(3)]))
    cat = dict(ndata=[1000003, 20], nrand=[5000000, 77], wdata=[1000003.25, 19.5],
               wrand=[4999999.5, 80.0], shot=[1234.5678901234, 0.25], norm=[0.031415926535, 2.0])
    meta = dict(Ng=256, bsize=[1000.0, 1250.5, 999.75], bmin=[-10.5, 0.0, 3.25], assign=2,
                intlace=1)
    return pk, cat, meta


def _call(pk, cat, meta, conf, files, header):
    keep = []

    def arr(a, t=C.c_double):
        a = np.ascontiguousarray(a, dtype=np.uint64 if t is C.c_uint64 else np.float64)
        keep.append(a)
        return a.ctypes.data

    p = P.Pk()
    p.nbin, p.nl = pk["nbin"], len(pk["poles"])
    for l, v in enumerate(pk["poles"]):
        p.poles[l] = v
    p.k, p.kedge, p.km = arr(pk["k"]), arr(pk["kedge"]), arr(pk["km"])
    p.cnt = arr(pk["cnt"], C.c_uint64)
    p.pl[0], p.pl[1] = arr(pk["pl"][0]), arr(pk["pl"][1])
    p.xpl = arr(pk["xpl"])
    ct = P.Cata()
    ct.num = 2
    fields = {}
    for key in ("ndata", "nrand"):
        fields[key] = (C.c_size_t * 2)(*cat[key])
    for key in ("wdata", "wrand", "alpha", "shot", "norm"):
        fields[key] = (C.c_double * 2)(*cat.get(key, [0.0, 0.0]))
    for key, v in fields.items():
        setattr(ct, key, C.cast(v, C.c_void_p))
    m = P.MeshMeta()
    m.Ng, m.assign, m.intlace = meta["Ng"], meta["assign"], meta["intlace"]
    m.bsize[:] = meta["bsize"]
    m.bmin[:] = meta["bmin"]
    fa = (C.c_char_p * 2)(*[f.encode() if f else None for f in files[:2]])
    fc = files[2].encode() if files[2] else None
    return P.load().psh_write_pk(C.byref(conf), C.byref(ct), C.byref(m), C.byref(p), fa, fc,
                                 int(header))


@pytest.mark.parametrize("issim", [True, False])
@pytest.mark.parametrize("header", [True, False])
def test_files_byte_exact(tmp_path, issim, header):
    pk, cat, meta = _case(issim, 5)
    conf = P.make_conf(gsize=256, bsize=1000.0, kbin=0.01, issim=issim, ndata=2,
                       isauto=[True, True], iscross=True, poles=pk["poles"])
    files = [str(tmp_path / "a0.txt"), str(tmp_path / "a1.txt"), str(tmp_path / "x.txt")]
    assert _call(pk, cat, meta, conf, files, header) == 0
    for n in range(2):
        exp = ""
        if header:
            exp += "# Data catalog: %d objects, total weight: %.10g\n" % (
                cat["ndata"][n], cat["wdata"][n])
            if not issim:
                exp += "# Random catalog: %d objects, total weight: %.10g\n" % (
                    cat["nrand"][n], cat["wrand"][n])
            exp += _render_mesh(meta)
            shot = cat["shot"][n] if issim else cat["shot"][n] / cat["norm"][n]
            exp += "# Shot noise: %.10g , normalisation: %.10g \n" % (shot, cat["norm"][n])
        exp += _render_cols(pk["nbin"], pk["poles"], pk["k"], pk["kedge"], pk["km"], pk["cnt"],
                            pk["pl"][n])
        assert open(files[n], "rb").read() == exp.encode()
    exp = ""
    if header:
        exp += "# Data catalog 1: %d objects, total weight: %.10g\n" % (cat["ndata"][0], cat["wdata"][0])
        exp += "# Data catalog 2: %d objects, total weight: %.10g\n" % (cat["ndata"][1], cat["wdata"][1])
        if not issim:
            exp += "# Random catalog 1: %d objects, total weight: %.10g\n" % (cat["nrand"][0], cat["wrand"][0])
            exp += "# Random catalog 2: %d objects, total weight: %.10g\n" % (cat["nrand"][1], cat["wrand"][1])
        exp += _render_mesh(meta)
        exp += "# Normalisation factor 1: %.10g , Normalisation factor 2: %.10g\n" % (
            cat["norm"][0], cat["norm"][1])
    exp += _render_cols(pk["nbin"], pk["poles"], pk["k"], pk["kedge"], pk["km"], pk["cnt"],
                        pk["xpl"])
    assert open(files[2], "rb").read() == exp.encode()


def test_only_requested_files(tmp_path):
    pk, cat, meta = _case(True, 6)
    conf = P.make_conf(gsize=256, bsize=1000.0, kbin=0.01, ndata=2, isauto=[False, True],
                       iscross=False, poles=pk["poles"])
    files = [None, str(tmp_path / "a1.txt"), None]
    assert _call(pk, cat, meta, conf, files, True) == 0
    assert os.listdir(tmp_path) == ["a1.txt"]


def test_errors(tmp_path):
    pk, cat, meta = _case(True, 7)
    conf = P.make_conf(gsize=256, bsize=1000.0, kbin=0.01, ndata=2, isauto=[True, False],
                       iscross=True, poles=pk["poles"])
    bad = [str(tmp_path / "nodir" / "a.txt"), None, str(tmp_path / "x.txt")]
    assert _call(pk, cat, meta, conf, bad, False) == P.ERR_FILE
    noname = [str(tmp_path / "a.txt"), None, None]   # cross requested without a name
    assert _call(pk, cat, meta, conf, noname, False) == P.ERR_ARG
