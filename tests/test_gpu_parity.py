"""GPU parity of the HIP path against the FP64 oracle, through the C-ABI.

Tolerances: mode counts bit-exact; for P_l per bin, every l,
    |P_l(gpu) - P_l(oracle)| <= TOL * S_b
with S_b = |P_0(oracle)| + shot (boxes) or |P_0| + shot/norm (surveys), the
un-subtracted monopole amplitude of the bin, and in bins whose |P_0| exceeds
the shot level also
    |P_0(gpu) - P_0(oracle)| <= TOL * |P_0(oracle)|,
with TOL = 1e-4 (FP32 meshes and FFTs against FP64).
"""
import math

import numpy as np
import pytest

TOL = 1e-4

pytestmark = pytest.mark.gpu


def _uniform(n, L, seed, weights=False):
    rng = np.random.default_rng(seed)
    d = np.empty((n, 4))
    d[:, :3] = rng.uniform(0.0, 1.0, (n, 3)) * np.asarray(L)
    d[:, 3] = rng.uniform(0.5, 1.5, n) if weights else 1.0
    return d


def _check(got, ref, poles, shot, norm=None, tol=TOL):
    assert got["nbin"] == ref["nbin"]
    assert np.array_equal(got["cnt"], ref["cnt"]), "mode counts differ"
    assert np.allclose(got["kedge"], ref["kedge"], rtol=0, atol=0)
    assert np.allclose(got["km"], ref["km"], rtol=1e-12)
    shot_level = shot if norm is None else shot / norm
    scale = np.abs(ref["pl"][0][0]) + shot_level
    has = ref["cnt"] > 0
    worst = 0.0
    for li, ell in enumerate(poles):
        d = np.abs(got["pl"][0][li] - ref["pl"][0][li])[has] / scale[has]
        worst = max(worst, float(d.max()) if d.size else 0.0)
        assert np.all(d <= tol), (ell, float(d.max()))
        assert np.all(got["pl"][0][li][~has] == ref["pl"][0][li][~has])
    if poles[0] == 0:
        p0 = ref["pl"][0][0]
        big = has & (np.abs(p0) > shot_level)
        rel = np.abs(got["pl"][0][0] - p0)[big] / np.abs(p0[big])
        assert np.all(rel <= tol), float(rel.max())
    return worst


def test_device_is_gfx950(psh):
    assert psh.device_arch().startswith("gfx950")


@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
@pytest.mark.parametrize("intlace", [False, True])
def test_sim_small_all_schemes(psh, oracle, scheme, intlace):
    Ng, L, n = 32, 1000.0, 30000
    d = _uniform(n, [L] * 3, 100 + scheme, weights=True)
    kw = dict(gsize=Ng, assign=scheme, intlace=intlace, bsize=L, kbin=0.01, poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw))
    got = ctx.run_device(psh.DeviceArray.from_numpy(d), n, d[:, 3].sum())
    ref = oracle.run(oracle.make_conf(**kw), [d])
    w = _check(got, ref, [0, 2, 4], ref["shot"][0])
    print(f"scheme={scheme} intlace={intlace} worst={w:.2e}")


@pytest.mark.parametrize("Ng", [8, 16, 64, 128])
def test_sim_grid_sizes(psh, oracle, Ng):
    L, n = 500.0, 20000
    d = _uniform(n, [L] * 3, Ng)
    kw = dict(gsize=Ng, assign=2, intlace=True, bsize=L, kbin=0.02, poles=[0, 1, 2, 3, 4, 6],
              los=[0.6, 0.0, 0.8])
    ctx = psh.Context(psh.make_conf(**kw))
    got = ctx.run_device(psh.DeviceArray.from_numpy(d), n, float(n))
    ref = oracle.run(oracle.make_conf(**kw), [d])
    _check(got, ref, [0, 1, 2, 3, 4, 6], ref["shot"][0])


@pytest.mark.parametrize("Ng,kmax", [(128, 0.3), (128, 0.6), (512, 0.3), (1024, 0.3)])
def test_sim_kmax_below_nyquist(psh, oracle, Ng, kmax):
    """k_max below k_Ny: most (k_y, k_z) tiles hold no binned mode; the z pass
    stores k_z only up to the last live tile, the y pass transforms and stores
    only what the x pass reads, and the x pass skips the rest (N_g = 512 and
    1024 take the staged and the line-by-line z pass)."""
    L, n = 500.0, 40000
    d = _uniform(n, [L] * 3, 57, weights=True)
    kw = dict(gsize=Ng, assign=3, intlace=True, bsize=L, kbin=0.01, kmax=kmax,
              poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw))
    got = ctx.run_device(psh.DeviceArray.from_numpy(d), n, d[:, 3].sum())
    ref = oracle.run(oracle.make_conf(**kw), [d])
    assert got["nbin"] == round(kmax / 0.01)
    _check(got, ref, [0, 2, 4], ref["shot"][0])


def test_sim_config1_exact_setup(psh, oracle):
    """L=1000, 1e6 uniform particles, CIC, Ng=256, no interlacing, l=0,2,4,
    dk=0.01 (the reference's CPU-runnable case)."""
    Ng, L, n = 256, 1000.0, 1_000_000
    d = _uniform(n, [L] * 3, 1)
    kw = dict(gsize=Ng, assign=1, intlace=False, bsize=L, kbin=0.01, poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw))
    got = ctx.run_device(psh.DeviceArray.from_numpy(d), n, float(n))
    ref = oracle.run(oracle.make_conf(**kw), [d])
    assert ref["nbin"] == 80
    w = _check(got, ref, [0, 2, 4], ref["shot"][0])
    print(f"config-1 worst={w:.2e}")


@pytest.mark.parametrize("Ng", [512, 1024])
def test_sim_shot_noise_large_grid(psh, oracle, Ng):
    """A shot-noise-only catalogue (2e6 uniform particles) on a large PCS
    interlaced grid: the lowest bins hold |F|^2 at the shot level, 1e6 times
    below the (unbinned) DC power, with PCS alias factors up to 37 per axis
    elsewhere on the grid.  Each x tile's fixed-point scale must follow the
    contributions it sums; scaled by the DC power and the grid's largest
    alias factors it quantised the low-k sums at 5e-4 of S_b."""
    L, n = 1000.0, 2_000_000
    d = _uniform(n, [L] * 3, 4242 + Ng)
    kw = dict(gsize=Ng, assign=psh.PCS, intlace=True, bsize=L, kbin=0.005, poles=[0, 2, 4])
    got = psh.Context(psh.make_conf(**kw)).run_device(psh.DeviceArray.from_numpy(d), n,
                                                      float(n))
    ref = oracle.run(oracle.make_conf(**kw), [d])
    w = _check(got, ref, [0, 2, 4], ref["shot"][0])
    print(f"shot noise Ng={Ng}: worst {w:.2e}")


def test_sim_clustered_pcs_interlaced(psh, oracle):
    """Clustered particles (many per cell, strong large-scale power)."""
    Ng, L = 128, 1000.0
    rng = np.random.default_rng(3)
    centres = rng.uniform(0, L, (400, 3))
    pts = (centres[rng.integers(0, 400, 300000)] + rng.normal(0, 15.0, (300000, 3))) % L
    d = np.empty((len(pts), 4))
    d[:, :3] = pts
    d[:, 3] = 1.0
    kw = dict(gsize=Ng, assign=3, intlace=True, bsize=L, kbin=0.005, poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw))
    got = ctx.run_device(psh.DeviceArray.from_numpy(d), len(d), float(len(d)))
    ref = oracle.run(oracle.make_conf(**kw), [d])
    w = _check(got, ref, [0, 2, 4], ref["shot"][0])
    print(f"clustered worst={w:.2e}")


def test_sim_boundary_particles(psh, oracle):
    """Particles on cell edges, on the last half cell (periodic shift of the
    interlaced grid) and at the largest coordinate below L."""
    Ng, L = 16, 160.0
    h = L / Ng
    pts = [[0, 0, 0], [h, 2 * h, 3 * h], [L - 1e-9, L - h / 2, L - h / 4],
           [h / 2, 1.5 * h, 15.5 * h], [np.nextafter(L, 0)] * 3, [7.5 * h, 8 * h, 0.25 * h]]
    d = np.array([p + [1.0] for p in pts] * 50)
    for scheme in range(4):
        kw = dict(gsize=Ng, assign=scheme, intlace=True, bsize=L, kbin=0.05, poles=[0, 2])
        ctx = psh.Context(psh.make_conf(**kw))
        got = ctx.run_device(psh.DeviceArray.from_numpy(d), len(d), float(len(d)))
        ref = oracle.run(oracle.make_conf(**kw), [d])
        _check(got, ref, [0, 2], ref["shot"][0])


def test_run_host_dropin_sim(psh, oracle):
    Ng, L, n = 32, 300.0, 5000
    d = _uniform(n, [L] * 3, 9, weights=True)
    kw = dict(gsize=Ng, assign=1, intlace=True, bsize=L, kbin=0.03, poles=[0, 2])
    ctx = psh.Context(psh.make_conf(**kw))
    got = ctx.run_host([d])
    ref = oracle.run(oracle.make_conf(**kw), [d])
    _check(got, ref, [0, 2], ref["shot"][0])
    assert math.isclose(got["shot"][0], ref["shot"][0], rel_tol=1e-15)
    assert math.isclose(got["norm"][0], ref["norm"][0], rel_tol=1e-15)


def test_box_errors(psh):
    """The box check of src/genr_mesh.c:517-533 (folded into the count pass)
    rejects x >= L and x < 0 with the reference's messages, in plain and
    batched runs, and leaves the context usable."""
    good = _uniform(100, [100.0] * 3, 2)
    ctx = psh.Context(psh.make_conf(gsize=16, bsize=100.0, kbin=0.05, intlace=True))
    for axis, val, msg in ((1, 100.0, "y coordinate not smaller than BOX_SIZE"),
                           (2, -1e-9, "z coordinate below 0")):
        d = good.copy()
        d[3, axis] = val
        with pytest.raises(psh.PshError) as e:
            ctx.run_device(psh.DeviceArray.from_numpy(d), len(d), 100.0)
        assert e.value.code == psh.ERR_CATA
        assert msg in str(e.value)
    ref = ctx.run_device(psh.DeviceArray.from_numpy(good), len(good), 100.0)
    assert np.isfinite(ref["pl"][0]).all()
    bctx = psh.Context(psh.make_conf(gsize=16, bsize=100.0, kbin=0.05), nmock_max=2)
    d = np.concatenate([good, good])
    d[150, 0] = 100.0
    with pytest.raises(psh.PshError) as e:
        bctx.run_batch(psh.DeviceArray.from_numpy(d), [0, 100, 200], [100.0, 100.0])
    assert e.value.code == psh.ERR_CATA


def test_batch_matches_plain(psh):
    Ng, L = 32, 1000.0
    rng = np.random.default_rng(11)
    sizes = rng.integers(2000, 9000, 6)
    cats = [_uniform(int(s), [L] * 3, 200 + i) for i, s in enumerate(sizes)]
    allp = np.concatenate(cats)
    off = np.concatenate([[0], np.cumsum(sizes)])
    kw = dict(gsize=Ng, assign=2, intlace=True, bsize=L, kbin=0.01, poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw), nmock_max=len(cats))
    dev = psh.DeviceArray.from_numpy(allp)
    batch = ctx.run_batch(dev, off, [float(s) for s in sizes], graph=False)
    graph = ctx.run_batch(dev, off, [float(s) for s in sizes], graph=True)
    graph2 = ctx.run_batch(dev, off, [float(s) for s in sizes], graph=True)
    single = psh.Context(psh.make_conf(**kw))
    for m, c in enumerate(cats):
        p = single.run_device(psh.DeviceArray.from_numpy(c), len(c), float(len(c)))
        for r in (batch[m], graph[m], graph2[m]):
            assert np.array_equal(r["cnt"], p["cnt"])
            scale = np.abs(p["pl"][0][0]) + 1e9 / len(c)
            assert np.all(np.abs(r["pl"][0] - p["pl"][0]) <= 1e-5 * scale)


def test_batch_graph_after_box_change(psh):
    """A captured batch graph is dropped when the box changes: the replay
    after psh_ctx_set_box uses the new box's tables, as a fresh context."""
    Ng = 32
    rng = np.random.default_rng(13)
    sizes = rng.integers(2000, 5000, 3)
    cats = [_uniform(int(v), [800.0] * 3, 400 + i) for i, v in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)])
    dev = psh.DeviceArray.from_numpy(np.concatenate(cats))
    w = [float(v) for v in sizes]
    kw = dict(gsize=Ng, assign=2, intlace=True, bsize=1000.0, kbin=0.01, poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw), nmock_max=len(cats))
    ctx.run_batch(dev, off, w, graph=True)
    ctx.set_box(800.0)
    got = ctx.run_batch(dev, off, w, graph=True)
    ref = psh.Context(psh.make_conf(**dict(kw, bsize=800.0)), nmock_max=len(cats))
    want = ref.run_batch(dev, off, w, graph=False)
    for m in range(len(cats)):
        assert np.array_equal(got[m]["cnt"], want[m]["cnt"])
        scale = np.abs(want[m]["pl"][0][0]) + 1e9 / sizes[m]
        assert np.all(np.abs(got[m]["pl"][0] - want[m]["pl"][0]) <= 1e-5 * scale)


def test_batch_graph_after_plain_run(psh):
    """A plain run between two graph replays may grow the partial-sum and
    bin buffers that the captured batch graph holds; the replay after it
    still matches eager runs (the graph is dropped on reallocation)."""
    Ng = 32
    rng = np.random.default_rng(14)
    sizes = rng.integers(2000, 5000, 3)
    cats = [_uniform(int(v), [1000.0] * 3, 500 + i) for i, v in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)])
    dev = psh.DeviceArray.from_numpy(np.concatenate(cats))
    w = [float(v) for v in sizes]
    kw = dict(gsize=Ng, assign=1, intlace=True, bsize=1000.0, kbin=0.01, poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw), nmock_max=len(cats))
    first = ctx.run_batch(dev, off, w, graph=True)
    one = ctx.run_device(psh.DeviceArray.from_numpy(cats[0]), len(cats[0]), w[0])
    again = ctx.run_batch(dev, off, w, graph=True)
    eager = ctx.run_batch(dev, off, w, graph=False)
    for m in range(len(cats)):
        for r in (first[m], again[m]):
            assert np.array_equal(r["cnt"], eager[m]["cnt"])
            scale = np.abs(eager[m]["pl"][0][0]) + 1e9 / sizes[m]
            assert np.all(np.abs(r["pl"][0] - eager[m]["pl"][0]) <= 1e-5 * scale)
    assert np.all(np.abs(one["pl"][0] - eager[0]["pl"][0])
                  <= 1e-5 * (np.abs(eager[0]["pl"][0][0]) + 1e9 / sizes[0]))


def test_batch_graph_new_offsets(psh):
    """Replaying a captured batch graph with the same data, mock count and
    total size but other per-mock offsets uses the new split."""
    Ng = 32
    allp = _uniform(12000, [1000.0] * 3, 600)
    dev = psh.DeviceArray.from_numpy(allp)
    kw = dict(gsize=Ng, assign=3, intlace=False, bsize=1000.0, kbin=0.01, poles=[0, 2])
    ctx = psh.Context(psh.make_conf(**kw), nmock_max=3)
    ref = psh.Context(psh.make_conf(**kw))   # eager runs elsewhere keep ctx's graph
    for off in ([0, 3000, 7000, 12000], [0, 5000, 6000, 12000]):
        w = [float(off[m + 1] - off[m]) for m in range(3)]
        got = ctx.run_batch(dev, np.asarray(off), w, graph=True)
        for m in range(3):
            one = psh.DeviceArray.from_numpy(allp[off[m]:off[m + 1]])
            want = ref.run_device(one, off[m + 1] - off[m], w[m])
            assert np.array_equal(got[m]["cnt"], want["cnt"])
            scale = np.abs(want["pl"][0][0]) + 1e9 / w[m]
            assert np.all(np.abs(got[m]["pl"][0] - want["pl"][0]) <= 1e-5 * scale)


def test_batch_streaming_matches_plain(psh):
    """A batch whose meshes outgrow the 256 MB Infinity Cache (16 interlaced
    mocks at N_g = 128: 335 MB of meshes) takes the non-temporal loads and
    stores; one mock alone (21 MB) takes the ordinary ones.  Both give the
    same spectra."""
    Ng, L = 128, 1000.0
    rng = np.random.default_rng(12)
    sizes = rng.integers(2000, 6000, 16)
    cats = [_uniform(int(s), [L] * 3, 300 + i) for i, s in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)])
    kw = dict(gsize=Ng, assign=2, intlace=True, bsize=L, kbin=0.01, poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw), nmock_max=len(cats))
    batch = ctx.run_batch(psh.DeviceArray.from_numpy(np.concatenate(cats)), off,
                          [float(v) for v in sizes], graph=False)
    single = psh.Context(psh.make_conf(**kw))
    for m in (0, 7, 15):
        c = cats[m]
        p = single.run_device(psh.DeviceArray.from_numpy(c), len(c), float(len(c)))
        assert np.array_equal(batch[m]["cnt"], p["cnt"])
        scale = np.abs(p["pl"][0][0]) + 1e9 / len(c)
        assert np.all(np.abs(batch[m]["pl"][0] - p["pl"][0]) <= 1e-5 * scale)


def _survey_cats(nrand, seed, nz=3e-4, p0=1e4):
    """Randoms uniform in a shell sector (r in [500, 900], RA in [0, 60) deg,
    Dec in [-20, 20) deg); data drawn from them with a modulated selection.
    Completeness weights in [0.8, 1.2], FKP weights 1/(1 + n p0)."""
    rng = np.random.default_rng(seed)
    r = np.cbrt(rng.uniform(500.0 ** 3, 900.0 ** 3, nrand))
    ra = np.radians(rng.uniform(0.0, 60.0, nrand))
    dec = np.arcsin(rng.uniform(np.sin(np.radians(-20.0)), np.sin(np.radians(20.0)), nrand))
    xyz = np.stack([r * np.cos(dec) * np.cos(ra), r * np.cos(dec) * np.sin(ra),
                    r * np.sin(dec)], axis=1)
    sel = 0.25 * (1 + 0.8 * np.sin(2 * np.pi * xyz[:, 0] / 210.0)
                  * np.cos(2 * np.pi * xyz[:, 2] / 170.0))
    dsel = rng.uniform(0, 1, nrand) < sel
    dat = xyz[dsel] + rng.normal(0, 3.0, (int(dsel.sum()), 3))
    out = []
    for pts in (dat, xyz):
        wc = rng.uniform(0.8, 1.2, len(pts))
        nzv = np.full(len(pts), nz)
        out.append((pts, wc, 1.0 / (1.0 + nzv * p0), nzv))
    return out


@pytest.mark.parametrize("scheme,intlace,Ng,poles", [
    (1, False, 64, [0, 2, 4]),
    (1, True, 64, [0, 2, 4]),
    (2, True, 32, [0, 1, 2, 3, 4]),
    (3, False, 32, [0, 2]),
    (0, True, 32, [0, 2, 4, 6]),
])
def test_survey_parity(psh, oracle, scheme, intlace, Ng, poles):
    (dp, dwc, dfkp, dnz), (rp, rwc, rfkp, rnz) = _survey_cats(120000, 40 + scheme)
    wd, wr, sums = oracle.survey_sums(dwc, dfkp, dnz, rwc, rfkp, rnz)
    d = np.column_stack([dp, wd])
    r = np.column_stack([rp, wr])
    kw = dict(gsize=Ng, assign=scheme, intlace=intlace, issim=False, kbin=0.01, poles=poles)
    ctx = psh.Context(psh.make_conf(**kw))
    got = ctx.run_host([d], [r], [sums])
    ref = oracle.run(oracle.make_conf(**kw), [d], [r], [sums])
    w = _check(got, ref, poles, sums["shot"], sums["norm"])
    print(f"survey scheme={scheme} intlace={intlace} Ng={Ng} worst={w:.2e}")


@pytest.mark.parametrize("cap", [1, 2, 4])
def test_survey_poles_split_over_groups(psh, oracle, monkeypatch, cap):
    """The (l, m) fields of a pole binned by several x passes when the field
    buffers hold fewer than 2l + 1 (PSH_SURVEY_FIELDS caps them): every
    split gives the oracle's multipoles, and one group's to FP32 rounding."""
    (dp, dwc, dfkp, dnz), (rp, rwc, rfkp, rnz) = _survey_cats(120000, 47)
    wd, wr, sums = oracle.survey_sums(dwc, dfkp, dnz, rwc, rfkp, rnz)
    d = np.column_stack([dp, wd])
    r = np.column_stack([rp, wr])
    kw = dict(gsize=32, assign=2, intlace=True, issim=False, kbin=0.01, poles=[0, 1, 2, 4])
    monkeypatch.delenv("PSH_SURVEY_FIELDS", raising=False)
    whole = psh.Context(psh.make_conf(**kw)).run_host([d], [r], [sums])
    monkeypatch.setenv("PSH_SURVEY_FIELDS", str(cap))
    got = psh.Context(psh.make_conf(**kw)).run_host([d], [r], [sums])
    ref = oracle.run(oracle.make_conf(**kw), [d], [r], [sums])
    _check(got, ref, [0, 1, 2, 4], sums["shot"], sums["norm"])
    # a group sums its fields per mode in FP32 before binning, so a split
    # differs from one group by FP32 rounding only
    for li in range(4):
        s = np.abs(whole["pl"][0][li]) + sums["shot"]
        assert np.all(np.abs(got["pl"][0][li] - whole["pl"][0][li]) <= 1e-5 * s)


def _check_cross(got, ref, poles, scale, tol=TOL):
    has = ref["cnt"] > 0
    worst = 0.0
    for li, ell in enumerate(poles):
        d = np.abs(got["xpl"][li] - ref["xpl"][li])[has] / scale[has]
        worst = max(worst, float(d.max()) if d.size else 0.0)
        assert np.all(d <= tol), (ell, float(d.max()))
    return worst


@pytest.mark.parametrize("scheme,intlace,isauto", [
    (1, False, [True, True]), (3, True, [True, False]), (2, True, [False, False])])
def test_cross_sim(psh, oracle, scheme, intlace, isauto):
    """Two correlated boxes: the second shares half of the first's particles."""
    Ng, L = 32, 800.0
    rng = np.random.default_rng(70 + scheme)
    c = rng.uniform(0, L, (60, 3))
    p0 = (c[rng.integers(0, 60, 40000)] + rng.normal(0, 25.0, (40000, 3))) % L
    p1 = np.concatenate([p0[:20000], rng.uniform(0, L, (15000, 3))])
    d0 = np.column_stack([p0, np.ones(len(p0))])
    d1 = np.column_stack([p1, rng.uniform(0.5, 1.5, len(p1))])
    kw = dict(gsize=Ng, assign=scheme, intlace=intlace, bsize=L, kbin=0.02, poles=[0, 2, 4],
              ndata=2, isauto=isauto, iscross=True)
    got = psh.Context(psh.make_conf(**kw)).run_host([d0, d1])
    ref = oracle.run(oracle.make_conf(**kw), [d0, d1])
    assert np.array_equal(got["cnt"], ref["cnt"])
    s = [np.abs(ref["pl"][c][0]) + ref["shot"][c] for c in range(2)]
    for c in range(2):
        if isauto[c]:
            _check({**got, "pl": [got["pl"][c]]}, {**ref, "pl": [ref["pl"][c]]}, [0, 2, 4],
                   ref["shot"][c])
    if not all(isauto):   # autos not returned: scale from a full oracle run
        full = oracle.run(oracle.make_conf(**{**kw, "isauto": [True, True]}), [d0, d1])
        s = [np.abs(full["pl"][c][0]) + full["shot"][c] for c in range(2)]
    w = _check_cross(got, ref, [0, 2, 4], np.sqrt(s[0] * s[1]))
    print(f"cross sim scheme={scheme} il={intlace} worst={w:.2e}")


@pytest.mark.parametrize("share", [0.9, 0.0])
def test_cross_sim_256(psh, oracle, share):
    """The cross spectrum from the sum field, (S[0+1] - S[0] - S[1]) / 2 with
    FP32 meshes, at N_g = 256 (TSC, interlaced): tracers that share 90 % of
    their objects, and independent ones whose cross power is near zero
    against their autos (the cancellation case), both within the bar of
    sqrt(S_0 S_1) per bin."""
    Ng, L, n = 256, 1000.0, 300000
    rng = np.random.default_rng(4242)
    c = rng.uniform(0, L, (400, 3))
    p0 = (c[rng.integers(0, 400, n)] + rng.normal(0, 20.0, (n, 3))) % L
    m = int(share * n)
    q = (c[rng.integers(0, 400, n - m)] + rng.normal(0, 20.0, (n - m, 3))) % L
    if share == 0.0:
        q = rng.uniform(0, L, (n, 3))
    p1 = np.concatenate([p0[:m], q])
    d0 = np.column_stack([p0, np.ones(n)])
    d1 = np.column_stack([p1, rng.uniform(0.5, 1.5, len(p1))])
    kw = dict(gsize=Ng, assign=psh.TSC, intlace=True, bsize=L, kbin=0.01, poles=[0, 2, 4],
              ndata=2, isauto=[True, True], iscross=True)
    got = psh.Context(psh.make_conf(**kw)).run_host([d0, d1])
    ref = oracle.run(oracle.make_conf(**kw), [d0, d1])
    assert np.array_equal(got["cnt"], ref["cnt"])
    s = [np.abs(ref["pl"][c][0]) + ref["shot"][c] for c in range(2)]
    w = _check_cross(got, ref, [0, 2, 4], np.sqrt(s[0] * s[1]))
    x0 = np.abs(ref["xpl"][0][ref["cnt"] > 0]) / np.sqrt(s[0] * s[1])[ref["cnt"] > 0]
    print(f"cross N_g=256 share={share}: worst {w:.2e}, median |P_x|/scale {np.median(x0):.2e}")


@pytest.mark.parametrize("scheme,intlace", [(1, True), (2, False)])
def test_cross_survey(psh, oracle, scheme, intlace):
    (dp, dwc, dfkp, dnz), (rp, rwc, rfkp, rnz) = _survey_cats(100000, 80 + scheme)
    wd, wr, s0 = oracle.survey_sums(dwc, dfkp, dnz, rwc, rfkp, rnz)
    d0 = np.column_stack([dp, wd])
    r0 = np.column_stack([rp, wr])
    # second tracer: half of the first's objects plus new ones, its own randoms
    (ep, ewc, efkp, enz), (qp, qwc, qfkp, qnz) = _survey_cats(80000, 90 + scheme)
    pts = np.concatenate([dp[: len(dp) // 2], ep])
    wc = np.concatenate([dwc[: len(dp) // 2], ewc])
    fk = np.concatenate([dfkp[: len(dp) // 2], efkp])
    nz = np.concatenate([dnz[: len(dp) // 2], enz])
    wd1, wr1, s1 = oracle.survey_sums(wc, fk, nz, qwc, qfkp, qnz)
    d1 = np.column_stack([pts, wd1])
    r1 = np.column_stack([qp, wr1])
    kw = dict(gsize=32, assign=scheme, intlace=intlace, issim=False, kbin=0.01,
              poles=[0, 2, 4], ndata=2, isauto=[True, True], iscross=True)
    got = psh.Context(psh.make_conf(**kw)).run_host([d0, d1], [r0, r1], [s0, s1])
    ref = oracle.run(oracle.make_conf(**kw), [d0, d1], [r0, r1], [s0, s1])
    for c, sm in enumerate((s0, s1)):
        _check({**got, "pl": [got["pl"][c]]}, {**ref, "pl": [ref["pl"][c]]}, [0, 2, 4],
               sm["shot"], sm["norm"])
    sc = [np.abs(ref["pl"][c][0]) + sm["shot"] / sm["norm"] for c, sm in enumerate((s0, s1))]
    w = _check_cross(got, ref, [0, 2, 4], np.sqrt(sc[0] * sc[1]))
    print(f"cross survey scheme={scheme} il={intlace} worst={w:.2e}")


def test_dropin_entry_points(psh, oracle):
    """psh_genr_mesh / psh_powspec / destructors with a CATA-layout struct whose
    particle arrays come from malloc (the library frees them, as the reference
    does), two boxes with the cross spectrum."""
    import ctypes as C
    libc = C.CDLL("libc.so.6")
    libc.malloc.restype = C.c_void_p
    libc.malloc.argtypes = [C.c_size_t]
    L = psh.load()
    L.psh_genr_mesh.restype = C.c_void_p
    L.psh_powspec.restype = C.POINTER(psh.Pk)
    L.psh_genr_mesh.argtypes = [C.c_void_p, C.c_void_p]
    L.psh_powspec.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.psh_mesh_destroy.argtypes = [C.c_void_p]
    L.psh_powspec_destroy.argtypes = [C.c_void_p]
    Ng, Lb = 32, 500.0
    cats = [_uniform(20000, [Lb] * 3, 300), _uniform(15000, [Lb] * 3, 301, weights=True)]
    cats[1][:5000, :3] = cats[0][:5000, :3]
    kw = dict(gsize=Ng, assign=2, intlace=True, bsize=Lb, kbin=0.03, poles=[0, 2],
              ndata=2, isauto=[True, True], iscross=True)
    conf = psh.make_conf(**kw)
    # the particle arrays and the two pointer arrays come from malloc: the
    # library frees all of them (src/genr_mesh.c:916-922)
    psz = C.sizeof(C.c_void_p)
    dptr_p = libc.malloc(2 * psz)
    rptr_p = libc.malloc(2 * psz)
    C.memset(rptr_p, 0, 2 * psz)
    dptr = (C.c_void_p * 2).from_address(dptr_p)
    nd = (C.c_size_t * 2)(*[len(c) for c in cats])
    vals = {k: (C.c_double * 2)() for k in ("wdata", "wrand", "alpha", "shot", "norm")}
    for i, c in enumerate(cats):
        p = libc.malloc(c.nbytes)
        C.memmove(p, c.ctypes.data, c.nbytes)
        dptr[i] = p
        vals["wdata"][i] = float(c[:, 3].sum())
    cat = psh.Cata()
    cat.num = 2
    cat.data = dptr_p
    cat.rand = rptr_p
    cat.ndata = C.cast(nd, C.c_void_p)
    cat.nrand = C.cast((C.c_size_t * 2)(), C.c_void_p)
    for k, v in vals.items():
        setattr(cat, k, C.cast(v, C.c_void_p))
    mesh = L.psh_genr_mesh(C.byref(conf), C.byref(cat))
    assert mesh
    assert not cat.data and not cat.rand              # freed like the reference
    pk = L.psh_powspec(C.byref(conf), C.byref(cat), mesh)
    assert pk
    nb, nl = pk.contents.nbin, pk.contents.nl
    arr = lambda p, n, t=C.c_double: np.ctypeslib.as_array(C.cast(p, C.POINTER(t)), (n,)).copy()
    got = {"nbin": nb, "cnt": arr(pk.contents.cnt, nb, C.c_uint64),
           "kedge": arr(pk.contents.kedge, nb + 1), "km": arr(pk.contents.km, nb),
           "pl": [arr(pk.contents.pl[c], nl * nb).reshape(nl, nb) for c in range(2)],
           "xpl": arr(pk.contents.xpl, nl * nb).reshape(nl, nb)}
    L.psh_powspec_destroy(pk)
    L.psh_mesh_destroy(C.c_void_p(mesh))
    ref = oracle.run(oracle.make_conf(**kw), cats)
    for c in range(2):
        assert math.isclose(vals["shot"][c], ref["shot"][c], rel_tol=1e-15)
        _check({**got, "pl": [got["pl"][c]]}, {**ref, "pl": [ref["pl"][c]]}, [0, 2],
               ref["shot"][c])
    s = [np.abs(ref["pl"][c][0]) + ref["shot"][c] for c in range(2)]
    _check_cross(got, ref, [0, 2], np.sqrt(s[0] * s[1]))


@pytest.mark.parametrize("issim", [True, False])
def test_log_bins(psh, oracle, issim):
    """LOG_SCALE binning (0.5 log10 |k|^2, src/multipole.c:147-158) with the
    reference's rescale of k, km and the edges (Q7, src/multipole.c:1267-1273)."""
    kw = dict(gsize=64, assign=2, intlace=True, logscale=True, kmin=math.log10(0.004),
              kmax=math.log10(0.5), kbin=0.05, poles=[0, 2, 4])
    if issim:
        d = _uniform(60000, [600.0] * 3, 17, weights=True)
        kw["bsize"] = 600.0
        got = psh.Context(psh.make_conf(**kw)).run_host([d])
        ref = oracle.run(oracle.make_conf(**kw), [d])
        shot, norm = ref["shot"][0], None
    else:
        (dp, dwc, dfkp, dnz), (rp, rwc, rfkp, rnz) = _survey_cats(100000, 21)
        wd, wr, sums = oracle.survey_sums(dwc, dfkp, dnz, rwc, rfkp, rnz)
        d = np.column_stack([dp, wd])
        r = np.column_stack([rp, wr])
        kw["issim"] = False
        got = psh.Context(psh.make_conf(**kw)).run_host([d], [r], [sums])
        ref = oracle.run(oracle.make_conf(**kw), [d], [r], [sums])
        shot, norm = sums["shot"], sums["norm"]
    assert ref["nbin"] > 10
    assert np.array_equal(got["k"], ref["k"])
    w = _check(got, ref, [0, 2, 4], shot, norm)
    print(f"log bins issim={issim} nbin={ref['nbin']} worst={w:.2e}")


def test_save_res_dropin(psh, tmp_path):
    """psh_save_res writes what psh_powspec returned, in the reference's format."""
    import ctypes as C
    d = _uniform(20000, [400.0] * 3, 23)
    kw = dict(gsize=32, assign=1, intlace=False, bsize=400.0, kbin=0.04, poles=[0, 2])
    ctx = psh.Context(psh.make_conf(**kw))
    res = ctx.run_host([d])
    L = psh.load()
    pk = psh.Pk()
    keep = [res["k"], res["kedge"], res["km"], res["cnt"], res["pl"][0].ravel().copy()]
    pk.nbin, pk.nl = res["nbin"], 2
    pk.poles[0], pk.poles[1] = 0, 2
    pk.k, pk.kedge, pk.km = (a.ctypes.data for a in keep[:3])
    pk.cnt = keep[3].ctypes.data
    pk.pl[0] = keep[4].ctypes.data
    meta = psh.MeshMeta()
    meta.Ng, meta.assign, meta.intlace = 32, 1, 0
    meta.bsize[:] = [400.0] * 3
    cat = psh.Cata()
    vals = {k: (C.c_double * 2)(res[k][0] if k in ("shot", "norm") else float(len(d)), 0.0)
            for k in ("wdata", "shot", "norm")}
    nd = (C.c_size_t * 2)(len(d), 0)
    cat.num = 1
    cat.ndata = C.cast(nd, C.c_void_p)
    for k, v in vals.items():
        setattr(cat, k, C.cast(v, C.c_void_p))
    conf = psh.make_conf(**kw)
    f = str(tmp_path / "pk.txt")
    fa = (C.c_char_p * 2)(f.encode(), None)
    assert L.psh_write_pk(C.byref(conf), C.byref(cat), C.byref(meta), C.byref(pk), fa, None, 1) == 0
    tab = np.loadtxt(f)
    assert np.allclose(tab[:, 0], res["k"], rtol=1e-9)
    assert np.array_equal(tab[:, 4].astype(np.uint64), res["cnt"])
    assert np.allclose(tab[:, 5], res["pl"][0][0], rtol=1e-9, atol=1e-300)


def test_survey_device_entry(psh, oracle):
    """psh_ctx_run_survey_device (catalogues in HBM) gives the oracle's survey
    spectra, like the host entry point."""
    (dp, dwc, dfkp, dnz), (rp, rwc, rfkp, rnz) = _survey_cats(100000, 61)
    wd, wr, sums = oracle.survey_sums(dwc, dfkp, dnz, rwc, rfkp, rnz)
    d = np.column_stack([dp, wd])
    r = np.column_stack([rp, wr])
    kw = dict(gsize=32, assign=2, intlace=True, issim=False, kbin=0.01, poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw))
    ctx.set_profiling(True)
    got = ctx.run_survey_device(psh.DeviceArray.from_numpy(d), len(d),
                                psh.DeviceArray.from_numpy(r), len(r),
                                sums["alpha"], sums["shot"], sums["norm"])
    ref = oracle.run(oracle.make_conf(**kw), [d], [r], [sums])
    _check(got, ref, [0, 2, 4], sums["shot"], sums["norm"])
    # the padded box of the catalogues (src/genr_mesh.c:427-578)
    side, lo = ctx.box()
    assert np.allclose(side, ref["bsize"], rtol=1e-12) and np.allclose(lo, ref["bmin"], rtol=1e-12)
    # stage times of a survey run: the l = 0 x pass, the l > 0 fields' x
    # passes and z + y passes, all inside the total
    ms = ctx.timings()
    assert all(v > 0 for v in (ms[2], ms[3], ms[5], ms[6]))
    assert ms[2] + ms[5] + ms[6] < ms[3]


def test_host_upload_over_staging_chunks(psh, oracle):
    """Host catalogues larger than the upload's pinned chunks (32 MiB, two
    reused) and not a multiple of them: the host entry point gives the same
    bits as the device entry point on the same records, a survey's data and
    randoms (two uploads through the same chunks) included."""
    n = 3 * (1 << 20) + 12345          # 96.4 MiB of records: 3 chunks + a part
    d = _uniform(n, [1000.0] * 3, 91, weights=True)
    kw = dict(gsize=64, assign=2, intlace=True, bsize=1000.0, kbin=0.02, poles=[0, 2])
    host = psh.Context(psh.make_conf(**kw)).run_host([d])
    dev = psh.Context(psh.make_conf(**kw)).run_device(psh.DeviceArray.from_numpy(d), n,
                                                      float(d[:, 3].sum()))
    assert np.array_equal(host["cnt"], dev["cnt"])
    for li in range(2):
        assert np.array_equal(host["pl"][0][li], dev["pl"][0][li])
    (dp, dwc, dfkp, dnz), (rp, rwc, rfkp, rnz) = _survey_cats(400000, 92)
    # randoms repeated to 1.6e6 records (51 MB) so that they span two chunks
    rp, rwc, rfkp, rnz = (np.concatenate([a] * 4) for a in (rp, rwc, rfkp, rnz))
    wd, wr, sums = oracle.survey_sums(dwc, dfkp, dnz, rwc, rfkp, rnz)
    sd, sr = np.column_stack([dp, wd]), np.column_stack([rp, wr])
    kw = dict(gsize=32, assign=1, intlace=True, issim=False, kbin=0.01, poles=[0, 2])
    h = psh.Context(psh.make_conf(**kw)).run_host([sd], [sr], [sums])
    g = psh.Context(psh.make_conf(**kw)).run_survey_device(
        psh.DeviceArray.from_numpy(sd), len(sd), psh.DeviceArray.from_numpy(sr), len(sr),
        sums["alpha"], sums["shot"], sums["norm"])
    assert np.array_equal(h["cnt"], g["cnt"])
    for li in range(2):
        s = np.abs(g["pl"][0][li]) + sums["shot"]
        # the box and weight sums come from the host scan on one side and
        # the device bounds pass on the other: equal to FP64 rounding
        assert np.all(np.abs(h["pl"][0][li] - g["pl"][0][li]) <= 1e-9 * s)


@pytest.mark.parametrize("intlace", [False, True])
def test_sim_ng512_streaming_parity(psh, oracle, intlace):
    """N_g = 512 (the learned-correction size): meshes of 537 MB per grid take
    the non-temporal path, and the z pass stages its output rows in LDS."""
    dev, n = psh.gen_clustered_box(ng=256, L=1000.0, nbar=0.01, f=0.8, seed=5)
    data = dev.to_numpy((n, 4))
    kw = dict(gsize=512, assign=1, intlace=intlace, bsize=1000.0, kbin=0.01, poles=[0, 2, 4])
    got = psh.Context(psh.make_conf(**kw)).run_device(dev, n, float(n))
    dev.free()
    ref = oracle.run(oracle.make_conf(**kw), [data])
    worst = _check(got, ref, [0, 2, 4], ref["shot"][0])
    print(f"Ng 512 intlace={intlace}: {n} particles, worst |dP_l|/S_b {worst:.2e}")


def test_survey_ng512_parity(psh, oracle):
    """Survey estimator at N_g = 512 (l = 0, 2: the Y_2m-weighted z passes on
    the staged, non-temporal path) against the oracle."""
    (dp, dwc, dfkp, dnz), (rp, rwc, rfkp, rnz) = _survey_cats(400000, 61)
    wd, wr, sums = oracle.survey_sums(dwc, dfkp, dnz, rwc, rfkp, rnz)
    d = np.column_stack([dp, wd])
    r = np.column_stack([rp, wr])
    kw = dict(gsize=512, assign=1, intlace=True, issim=False, kbin=0.01, poles=[0, 2])
    ctx = psh.Context(psh.make_conf(**kw))
    got = ctx.run_survey_device(psh.DeviceArray.from_numpy(d), len(d),
                                psh.DeviceArray.from_numpy(r), len(r),
                                sums["alpha"], sums["shot"], sums["norm"])
    ref = oracle.run(oracle.make_conf(**kw), [d], [r], [sums])
    _check(got, ref, [0, 2], sums["shot"], sums["norm"])


def test_box1024_full_size_parity(psh, oracle):
    """The clustered-box workload at its full size (2.0e7 clustered particles, PCS,
    N_g = 1024, interlaced, l = 0,2,4, dk = 0.005) against the FP64 oracle:
    mode counts bit-exact, every P_l within the bar of the smaller tests."""
    dev, n = psh.gen_clustered_box(ng=512, L=1000.0, nbar=0.02, f=0.8, seed=2)
    data = dev.to_numpy((n, 4))
    kw = dict(gsize=1024, assign=3, intlace=True, bsize=1000.0, kbin=0.005, poles=[0, 2, 4])
    got = psh.Context(psh.make_conf(**kw)).run_device(dev, n, float(n))
    dev.free()
    ref = oracle.run(oracle.make_conf(**kw), [data])
    worst = _check(got, ref, [0, 2, 4], ref["shot"][0])
    print(f"box N_g=1024 full size: {n} particles, {len(ref['cnt'])} bins, worst |dP_l|/S_b {worst:.2e}")
