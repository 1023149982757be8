"""Catalogue ingest on the GPU: binary files and host columns -> device
psh_data, with the weight rule and catalogue sums of io/read_ascii.c:875-903."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cols(n, seed, survey=True):
    rng = np.random.default_rng(seed)
    c = {"x": rng.uniform(0, 1000, n), "y": rng.uniform(0, 1000, n),
         "z": rng.uniform(0, 1000, n), "w_comp": rng.uniform(0.5, 1.5, n)}
    if survey:
        c["w_fkp"] = rng.uniform(0.1, 1.0, n)
        c["nz"] = rng.uniform(1e-4, 3e-4, n)
    return c


def _expect(cols, dtype, sim):
    """psh_data rows and (sum w_comp, sum w^2, sum w_comp w_fkp^2 nz) with
    exactly rounded sums (math.fsum)."""
    c = {k: np.asarray(v, dtype).astype(np.float64) for k, v in cols.items()}
    n = len(c["x"])
    wc = c.get("w_comp", np.ones(n))
    if sim:
        w = wc
        s = (math.fsum(wc), 0.0, 0.0)
    else:
        wf = c.get("w_fkp", np.ones(n))
        nz = c.get("nz", np.zeros(n))
        w = wc * wf
        s = (math.fsum(wc), math.fsum(w * w), math.fsum(wc * wf * wf * nz))
    return np.stack([c["x"], c["y"], c["z"], w], 1), s


def _check(psh, got, res, cols, dtype, sim):
    rows, s = _expect(cols, dtype, sim)
    assert res["n"] == len(rows)
    assert np.array_equal(got, rows)        # copies and one product: bit-exact
    for a, b in zip((res["sumw"], res["sumw2"], res["sumw2n"]), s):
        assert abs(a - b) <= 1e-13 * abs(b) + 1e-300


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("sim", [True, False])
def test_file_ingest_matches(psh, tmp_path, dtype, sim):
    n = 300_001                              # ragged: not a multiple of the chunk
    cols = _cols(n, 1 + sim, survey=not sim)
    path = tmp_path / "cat.pscat"
    psh.cat_write(path, cols, dtype)
    out = psh.DeviceArray.empty((n + 10, 4), np.float64)
    ing = psh.Ingest(chunk_objects=65536)
    ing.start_file(path, out, n + 10, sim=sim)
    res = ing.wait()
    _check(psh, out.to_numpy((n + 10, 4), np.float64)[:n], res, cols, dtype, sim)
    ing.close()


def test_host_ingest_and_defaults(psh):
    n = 70_000
    cols = _cols(n, 5)
    del cols["w_comp"], cols["nz"]           # absent columns: w_comp = 1, nz = 0
    out = psh.DeviceArray.empty((n, 4), np.float64)
    ing = psh.Ingest(chunk_objects=30_000)
    ing.start_host(cols, out, n, sim=False)
    res = ing.wait()
    _check(psh, out.to_numpy((n, 4), np.float64), res, cols, np.float64, False)
    assert res["sumw"] == n and res["sumw2n"] == 0.0


def test_ingest_errors(psh, tmp_path):
    ing = psh.Ingest(chunk_objects=4096)
    out = psh.DeviceArray.empty((100, 4), np.float64)
    cols = _cols(200, 7)
    path = tmp_path / "c.pscat"
    psh.cat_write(path, cols)
    with pytest.raises(psh.PshError) as e:       # capacity
        ing.start_file(path, out, 100)
    assert e.value.code == psh.ERR_ARG
    raw = path.read_bytes()
    (tmp_path / "t.pscat").write_bytes(raw[:-8])  # truncated
    with pytest.raises(psh.PshError) as e:
        ing.start_file(tmp_path / "t.pscat", out, 1000)
    assert e.value.code == psh.ERR_FILE
    (tmp_path / "m.pscat").write_bytes(b"NOTACAT!" + raw[8:])
    with pytest.raises(psh.PshError) as e:
        ing.start_file(tmp_path / "m.pscat", out, 1000)
    assert e.value.code == psh.ERR_FILE
    with pytest.raises(psh.PshError):
        ing.start_file(tmp_path / "absent.pscat", out, 1000)
    # the object is still usable
    out2 = psh.DeviceArray.empty((200, 4), np.float64)
    ing.start_file(path, out2, 200, sim=False)
    res = ing.wait()
    _check(psh, out2.to_numpy((200, 4), np.float64), res, cols, np.float64, False)


def test_ingest_overlaps_processing(psh, tmp_path):
    """Catalogue k+1 is ingested on the ingest's stream while the context
    processes catalogue k; every P_l equals a run on the same data uploaded
    directly."""
    from powspec_amd import synth
    L, Ng, n = 500.0, 64, 200_000
    kw = dict(gsize=Ng, assign=2, intlace=True, bsize=L, kbin=0.02, poles=[0, 2])
    ctx = psh.Context(psh.make_conf(**kw))
    cats = [synth.uniform_box(n, L, seed=40 + k) for k in range(3)]
    paths = []
    for k, d in enumerate(cats):
        paths.append(tmp_path / f"c{k}.pscat")
        psh.cat_write(paths[-1], {"x": d[:, 0], "y": d[:, 1], "z": d[:, 2]})
    bufs = [psh.DeviceArray.empty((n, 4), np.float64) for _ in range(2)]
    ing = psh.Ingest(chunk_objects=1 << 16)
    ing.start_file(paths[0], bufs[0], n)
    res = ing.wait()
    got = []
    for k in range(3):
        if k + 1 < 3:
            ing.start_file(paths[k + 1], bufs[(k + 1) % 2], n)
        got.append(ctx.run_device(bufs[k % 2], n, res["sumw"]))
        if k + 1 < 3:
            res = ing.wait()
    for k in range(3):
        ref = ctx.run_device(psh.DeviceArray.from_numpy(cats[k]), n, float(n))
        for l in range(2):   # FP64 bin sums: equal up to their summation order
            a, b = got[k]["pl"][0][l], ref["pl"][0][l]
            assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())


def _ascii_catalogue(n, seed):
    """Lines in the reference's ASCII layout with varied number spellings:
    RA Dec z name w_comp w_fkp nz, plus a header, comments, blank lines,
    tabs and CRLF endings.  Returns (text, expected columns, header lines)."""
    rng = np.random.default_rng(seed)
    spell = [lambda v: repr(float(v)), lambda v: f"{v:.10g}", lambda v: f"{v:.6e}",
             lambda v: f"{v:+.12f}", lambda v: f"{v:.17E}"]
    header = ["# ra dec z name wc wfkp nz", "some header text"]
    lines = list(header)
    rows = []
    for i in range(n):
        vals = [rng.uniform(0, 360), rng.uniform(-90, 90), rng.uniform(0.1, 1.2),
                rng.uniform(0.5, 1.5), rng.uniform(0.05, 1.0), rng.uniform(1e-5, 5e-4)]
        toks = [spell[int(rng.integers(0, len(spell)))](v) for v in vals]
        rows.append([float(t) for t in toks])      # correctly rounded, as strtod
        sep = "\t" if i % 7 == 0 else " "
        ln = sep.join(toks[:3] + [f"gal{i}"] + toks[3:])
        if i % 11 == 0:
            ln = "   " + ln + "   "
        lines.append(ln + ("\r" if i % 13 == 0 else ""))
        if i % 17 == 0:
            lines.append("# a comment line")
        if i % 19 == 0:
            lines.append("   ")
    return "\n".join(lines), np.array(rows), len(header)


@pytest.mark.parametrize("chunk", [1000, 1 << 20])
def test_ascii_ingest_matches(psh, tmp_path, chunk):
    n = 20_003
    text, rows, nhead = _ascii_catalogue(n, 21)
    path = tmp_path / "cat.txt"
    path.write_text(text)                     # no newline after the last line
    out = psh.DeviceArray.empty((n, 4))
    ing = psh.Ingest(chunk_objects=chunk)     # 1000 objects: 48 KB chunks
    ing.start_ascii(path, out, n, fmtr="%lf %lf %lf %*s %lf %lf %lf",
                    pos=("$1", "$2", "${3}"), wcomp="$4", wfkp="$5", nz="$6",
                    skip=nhead, comment="#", sim=False)
    res = ing.wait()
    cols = {"x": rows[:, 0], "y": rows[:, 1], "z": rows[:, 2], "w_comp": rows[:, 3],
            "w_fkp": rows[:, 4], "nz": rows[:, 5]}
    _check(psh, out.to_numpy((n, 4)), res, cols, np.float64, False)


def test_ascii_matches_binary(psh, tmp_path):
    from powspec_amd import synth
    n = 50_000
    d = synth.uniform_box(n, 1000.0, seed=5)
    path = tmp_path / "box.txt"
    path.write_text("\n".join(f"{float(x)!r} {y:.17g} {z:.17e}" for x, y, z in d[:, :3])
                    + "\n")
    a = psh.DeviceArray.empty((n, 4))
    ing = psh.Ingest()
    ing.start_ascii(path, a, n, comment=None)
    ra = ing.wait()
    assert np.array_equal(a.to_numpy((n, 4)), d)
    assert ra["n"] == n and ra["sumw"] == n


def test_ascii_errors(psh, tmp_path):
    ing = psh.Ingest(chunk_objects=1000)
    out = psh.DeviceArray.empty((100, 4))
    lines = [f"{i}.5 2 3" for i in range(60)]
    lines[41] = "1.5 2 abc"
    p = tmp_path / "bad.txt"
    p.write_text("# head\n" + "\n".join(lines) + "\n")
    ing.start_ascii(p, out, 100, skip=1)
    with pytest.raises(psh.PshError) as e:
        ing.wait()
    assert e.value.code == psh.ERR_FILE and "line 43 " in str(e.value)
    lines[41] = "1.5 2"                          # missing column
    p.write_text("\n".join(lines) + "\n")
    ing.start_ascii(p, out, 100)
    with pytest.raises(psh.PshError) as e:
        ing.wait()
    assert "line 42 " in str(e.value)
    lines[41] = "1.5 2 3"
    p.write_text("\n".join(lines) + "\n")
    ing.start_ascii(p, out, 50)                   # capacity
    with pytest.raises(psh.PshError) as e:
        ing.wait()
    assert e.value.code == psh.ERR_ARG
    with pytest.raises(psh.PshError):             # unsupported formatter
        ing.start_ascii(p, out, 100, fmtr="%f %f %f")
    with pytest.raises(psh.PshError):             # column beyond the formatter
        ing.start_ascii(p, out, 100, pos=(1, 2, 4))
    ing.start_ascii(p, out, 100)
    r = ing.wait()
    assert r["n"] == 60
