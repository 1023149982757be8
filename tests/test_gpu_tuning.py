"""Deposit tile tuning: every tile shape gives the oracle's spectra (to the
same tolerance as the default), the autotuner picks a valid shape and its
cache entry is found again by psh_ctx_load_tuning."""
import numpy as np
import pytest

from test_gpu_parity import _check, _uniform

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tile", [(4, 4, 128), (8, 8, 64), (16, 16, 16), (32, 16, 16),
                                  (4, 32, 64), (4, 4, 8)])
def test_tile_shapes_keep_parity(psh, oracle, tile):
    Ng, L, n = 64, 500.0, 60000
    d = _uniform(n, [L] * 3, 51, weights=True)
    kw = dict(gsize=Ng, assign=3, intlace=True, bsize=L, kbin=0.02, poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw))
    ctx.set_tiles(tile)
    assert ctx.tiles() == tuple(min(t, Ng) for t in tile)   # the shape is the one in use
    got = ctx.run_device(psh.DeviceArray.from_numpy(d), n, d[:, 3].sum())
    ref = oracle.run(oracle.make_conf(**kw), [d])
    _check(got, ref, [0, 2, 4], ref["shot"][0])


def test_invalid_tiles_rejected(psh):
    ctx = psh.Context(psh.make_conf(gsize=64, bsize=500.0, kbin=0.02, intlace=True))
    for bad in [(3, 8, 32), (8, 8, 2), (64, 64, 64)]:
        with pytest.raises(psh.PshError) as e:
            ctx.set_tiles(bad)
        assert e.value.code == psh.ERR_ARG
    # PCS: tiles narrower than the 4-cell window, or whose padded z rows
    # (4 cells each side) overflow the LDS
    pcs = psh.Context(psh.make_conf(gsize=64, assign=3, bsize=500.0, kbin=0.02, intlace=True))
    for bad in [(2, 8, 32), (8, 2, 32), (32, 32, 16)]:
        with pytest.raises(psh.PshError) as e:
            pcs.set_tiles(bad)
        assert e.value.code == psh.ERR_ARG


def test_many_bins_in_windows(psh, oracle):
    """Binnings whose sums exceed the fused x pass's LDS are binned in
    windows of bins, one launch each (the reference has no bin limit,
    src/multipole.c:340-356): 8042 bins x 3 poles at N_g = 256 for a box and
    over 10^4 bins of a survey at N_g = 128, against the oracle."""
    from test_gpu_parity import _survey_cats
    d = _uniform(200000, [1000.0] * 3, 54)
    kw = dict(gsize=256, assign=2, bsize=1000.0, kbin=1e-4, intlace=True, poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw))
    assert ctx.nbin() == 8042
    got = ctx.run_device(psh.DeviceArray.from_numpy(d), len(d), float(len(d)))
    ref = oracle.run(oracle.make_conf(**kw), [d])
    _check(got, ref, [0, 2, 4], ref["shot"][0])
    (dp, dwc, dfkp, dnz), (rp, rwc, rfkp, rnz) = _survey_cats(100000, 55)
    wd, wr, sums = oracle.survey_sums(dwc, dfkp, dnz, rwc, rfkp, rnz)
    sd, sr = np.column_stack([dp, wd]), np.column_stack([rp, wr])
    kw = dict(gsize=128, assign=1, intlace=True, issim=False, kbin=2e-5, poles=[0, 2])
    got = psh.Context(psh.make_conf(**kw)).run_host([sd], [sr], [sums])
    ref = oracle.run(oracle.make_conf(**kw), [sd], [sr], [sums])
    assert ref["nbin"] > 10000
    _check(got, ref, [0, 2], sums["shot"], sums["norm"])


def test_autotune_and_cache(psh, tmp_path):
    cache = tmp_path / "tune.txt"
    conf = psh.make_conf(gsize=64, assign=2, intlace=True, bsize=500.0, kbin=0.02)
    tile, ms, ncand = psh.autotune(conf, 100000, 1, str(cache))
    d0 = psh.DeviceArray.from_numpy(_uniform(50000, [500.0] * 3, 53))
    t2, ms2, n2 = psh.autotune(conf, nmock=1, data=d0, offsets=[0, 50000])
    assert n2 == ncand and ms2 > 0
    assert ncand >= 8 and ms > 0
    assert 2048 <= tile[0] * tile[1] * tile[2] <= 16384
    line = cache.read_text().split()
    assert line[0].startswith("gfx950") and [int(v) for v in line[6:9]] == list(tile)
    ctx = psh.Context(conf)
    ctx.load_tuning(str(cache))
    assert ctx.tiles() == tuple(min(t, 64) for t in tile)
    d = _uniform(20000, [500.0] * 3, 52)
    res = ctx.run_device(psh.DeviceArray.from_numpy(d), len(d), float(len(d)))
    assert res["nbin"] > 0
    other = psh.Context(psh.make_conf(gsize=32, assign=2, intlace=True, bsize=500.0, kbin=0.02))
    with pytest.raises(psh.PshError):
        other.load_tuning(str(cache))     # no entry for Ng = 32


@pytest.mark.parametrize("Ng", [64, 256, 1024])
def test_fft_schedules_keep_parity(psh, oracle, Ng):
    """Every y-pass schedule that applies at a grid size (columns per tile,
    values per thread, radix plan) gives the oracle's spectra; the tuner
    times them, writes its choice to the cache and a context applies it."""
    L = 1000.0
    d = _uniform(150000 if Ng < 1024 else 400000, [L] * 3, 60 + Ng, weights=True)
    dev = psh.DeviceArray.from_numpy(d)
    kw = dict(gsize=Ng, assign=3, intlace=True, bsize=L, kbin=0.01 if Ng < 1024 else 0.02,
              poles=[0, 2, 4])
    ref = oracle.run(oracle.make_conf(**kw), [d])
    ctx = psh.Context(psh.make_conf(**kw))
    ran = []
    for ys in range(6):
        try:
            ctx.set_fft_schedule(ys)
        except psh.PshError as e:
            assert e.code == psh.ERR_ARG
            continue
        assert ctx.fft_schedule() == ys
        got = ctx.run_device(dev, len(d), float(d[:, 3].sum()))
        _check(got, ref, [0, 2, 4], ref["shot"][0])
        ran.append(ys)
    assert 0 in ran and len(ran) >= 3
    import tempfile, os
    with tempfile.TemporaryDirectory() as tmp:
        cache = os.path.join(tmp, "tune.txt")
        ys, ms, ncand, ms0 = psh.autotune_fft(psh.make_conf(**kw), 1, cache)
        assert ncand == len(ran) and ys in ran and 0 < ms <= ms0
        other = psh.Context(psh.make_conf(**kw))
        other.load_tuning(cache)
        assert other.fft_schedule() == ys
    print(f"Ng {Ng}: schedules {ran}, tuned {ys} ({ms:.3f} ms vs {ms0:.3f} ms default)")
