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
