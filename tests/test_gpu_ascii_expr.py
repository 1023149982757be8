"""ASCII ingest with libast expressions (psh_ingest_ascii_expr) against the
values the reference's own libast gives on the same rows
(tests/golden/ascii_expr_vectors.json, written by
tests/golden/make_ascii_expr_vectors.py from lib/libast.c).

Bar: selections bit-exact (the same rows kept, in order); values bit-exact
except where the device's pow / log / log10 round differently from the
host C library's, within 2 ulp."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

V = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                "ascii_expr_vectors.json")))
ROWS = V["rows"]
N = len(ROWS)
COL = np.array([[float(x) for x in r] for r in ROWS])   # value columns $1..$6


def _file(tmp_path):
    p = tmp_path / "expr.txt"
    p.write_text("# a b c d e f\n" + "\n".join(" ".join(r) for r in ROWS) + "\n")
    return p


def _run(psh, path, chunk=0, **kw):
    out = psh.DeviceArray.empty((N, 4))
    ing = psh.Ingest(chunk_objects=chunk)
    try:
        ing.start_ascii_expr(path, out, N, fmtr=V["formatter"], **kw)
        res = ing.wait()
        return out.to_numpy((N, 4))[:res["n"]], res
    finally:
        ing.close()
        out.free()


def _close(got, exp):
    exp = np.asarray(exp, dtype=np.float64)
    fin = np.isfinite(exp)
    assert np.array_equal(got[~fin], exp[~fin], equal_nan=True)   # inf, nan
    d = np.abs(got[fin] - exp[fin])
    assert np.all(d <= 2 * np.spacing(np.abs(exp[fin]))), float(d.max())
    return int(np.sum(d == 0)), int(np.sum(fin))


def _cases(kind, builds=True):
    return [c for c in V["cases"] if c["kind"] == kind and c["builds"] == builds]


def test_value_expressions(psh, tmp_path):
    p = _file(tmp_path)
    for c in _cases("D"):
        got, res = _run(psh, p, pos=(c["expr"], "$1", "$3"), wcomp=c["expr"], sim=True)
        assert res["n"] == N
        ex, tot = _close(got[:, 0], c["values"])
        _close(got[:, 3], c["values"])
        assert np.array_equal(got[:, 1], COL[:, 0]) and np.array_equal(got[:, 2], COL[:, 2])
        print(f"{c['expr']!r}: {ex}/{tot} bit-exact")


def test_selections(psh, tmp_path):
    p = _file(tmp_path)
    for c in _cases("B"):
        if c["eval_error"]:
            continue
        keep = np.array(c["keep"])
        got, res = _run(psh, p, pos=("$1", "$3", "$4"), sel=c["expr"])
        assert res["n"] == keep.sum(), c["expr"]
        assert np.array_equal(got[:, :3], COL[keep][:, [0, 2, 3]]), c["expr"]


def test_weights_sums_and_chunks(psh, tmp_path):
    """Survey weights from expressions, a selection and 50-object chunks:
    w = w_comp w_fkp and the three catalogue sums (io/read_ascii.c:875-903)
    from libast's values of the kept rows."""
    by = {(c["kind"], c["expr"]): c for c in V["cases"]}
    wc = np.array(by["D", "abs($4)+0.5"]["values"])
    wf = np.array(by["D", "1/(1+$6*$2)"]["values"])
    nz = np.array(by["D", "$6*1e-4"]["values"])
    keep = np.array(by["B", "$1 > 0"]["keep"])
    got, res = _run(psh, _file(tmp_path), chunk=50, pos=("$1", "${3}", "$4"),
                    wcomp="abs($4)+0.5", wfkp="1/(1+$6*$2)", nz="$6*1e-4", sel="$1 > 0",
                    sim=False)
    assert res["n"] == keep.sum()
    w = wc[keep] * wf[keep]
    assert np.array_equal(got[:, 3], w)
    assert res["sumw"] == pytest.approx(wc[keep].sum(), rel=1e-14)
    assert res["sumw2"] == pytest.approx((w * w).sum(), rel=1e-14)
    assert res["sumw2n"] == pytest.approx((wc[keep] * wf[keep] ** 2 * nz[keep]).sum(),
                                          rel=1e-14)


def test_refused(psh, tmp_path):
    """Expressions libast does not build or cannot evaluate: PSH_ERR_ARG at
    start."""
    p = _file(tmp_path)
    out = psh.DeviceArray.empty((N, 4))
    ing = psh.Ingest()
    try:
        bad_d = [c["expr"] for c in _cases("D", builds=False)]
        bad_b = [c["expr"] for c in _cases("B", builds=False)]
        bad_b += [c["expr"] for c in _cases("B") if c["eval_error"]] + V["refused_here"]
        for e in bad_d:
            with pytest.raises(psh.PshError) as ex:
                ing.start_ascii_expr(p, out, N, fmtr=V["formatter"], pos=(e, "$1", "$3"))
            assert ex.value.code == psh.ERR_ARG, e
        for e in bad_b:
            with pytest.raises(psh.PshError) as ex:
                ing.start_ascii_expr(p, out, N, fmtr=V["formatter"], sel=e)
            assert ex.value.code == psh.ERR_ARG, e
        with pytest.raises(psh.PshError):      # a position expression missing
            ing.start_ascii_expr(p, out, N, fmtr=V["formatter"], pos=("$1", None, "$3"))
        # still usable after the refusals
        ing.start_ascii_expr(p, out, N, fmtr=V["formatter"])
        assert ing.wait()["n"] == N
    finally:
        ing.close()
        out.free()


def test_braced_columns_and_skipped_fields(psh, tmp_path):
    """${n} past 9 and value columns counted among the conversions that are
    not suppressed: a 13-field formatter with %*s and %*lf fields, the
    positions read from ${10}, ${11} and $2 with a selection on ${12}."""
    rng = np.random.default_rng(7)
    n = 500
    vals = rng.uniform(0.0, 1.0, (n, 12))
    lines = []
    for r in vals:
        t = [repr(float(x)) for x in r]
        lines.append(" ".join(t[:3] + ["tag"] + t[3:6] + ["9.75"] + t[6:]))
    p = tmp_path / "wide.txt"
    p.write_text("\n".join(lines) + "\n")
    fmtr = "%lf %lf %lf %*s %lf %lf %lf %*lf %lf %lf %lf %lf %lf %lf"
    out = psh.DeviceArray.empty((n, 4))
    ing = psh.Ingest()
    try:
        ing.start_ascii_expr(p, out, n, fmtr=fmtr, pos=("${10}", "${11}", "$2"),
                             wcomp="${12}", sel="${12} >= 0.25")
        res = ing.wait()
        keep = vals[:, 11] >= 0.25
        got = out.to_numpy((n, 4))[:res["n"]]
        assert res["n"] == keep.sum()
        assert np.array_equal(got, vals[keep][:, [9, 10, 1, 11]])
        assert res["sumw"] == pytest.approx(vals[keep, 11].sum(), rel=1e-14)
    finally:
        ing.close()
        out.free()
