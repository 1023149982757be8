"""The z transform fused into the deposit (tiles of whole z rows): against
the FP64 oracle, against the separate deposit + z pass on the same input,
in batched graph replays, with k_z pruning (k_max below Nyquist), and
bit for bit from run to run.

Bar as in test_gpu_parity: counts bit-exact, |dP_l| <= 1e-4 S_b per bin; the
fused path's worst error is at most a few times the separate path's (the
same FFT passes; the tiles differ in shape, and so in fixed-point step)."""
import numpy as np
import pytest

from tests.test_gpu_parity import _check, _uniform

pytestmark = pytest.mark.gpu


def test_fused_tile_shapes(psh):
    for Ng in (128, 256, 512, 1024):
        for il in (False, True):
            t = psh.fused_tile(Ng, il)
            assert t is not None and t[2] == Ng and min(t[0], t[1]) >= 4
            # 32768 / Ng lines of Ng / 2 points over both grids (16384 / Ng
            # with the 64-bit cells of N_g <= 256)
            lines = (16384 if Ng <= 256 else 32768) // Ng
            assert t[0] * t[1] * (2 if il else 1) == lines
    assert psh.fused_tile(64, True) is None and psh.fused_tile(768, True) is None


def _run(psh, kw, d, tile):
    ctx = psh.Context(psh.make_conf(**kw))
    if tile is not None:
        ctx.set_tiles(tile)
    res = ctx.run_device(psh.DeviceArray.from_numpy(d), len(d), float(d[:, 3].sum()))
    return ctx, res


@pytest.mark.parametrize("Ng", [128, 256, 512])
@pytest.mark.parametrize("scheme,intlace", [(0, False), (1, True), (2, False), (3, True)])
def test_fused_vs_oracle(psh, oracle, Ng, scheme, intlace):
    L, n = 800.0, 200000
    d = _uniform(n, [L] * 3, 700 + Ng + 10 * scheme, weights=True)
    kw = dict(gsize=Ng, assign=scheme, intlace=intlace, bsize=L, kbin=0.01,
              poles=[0, 2, 4], los=[0.3, 0.4, np.sqrt(0.75)])
    tile = psh.fused_tile(Ng, intlace)
    ctx, got = _run(psh, kw, d, tile)
    assert ctx.tiles() == tile
    ref = oracle.run(oracle.make_conf(**kw), [d])
    w = _check(got, ref, [0, 2, 4], ref["shot"][0])
    _, sep = _run(psh, kw, d, (8, 16, 64))
    ws = _check(sep, ref, [0, 2, 4], ref["shot"][0])
    # no worse than the separate path beyond the fixed-point step of its
    # larger tiles (both are far inside the bar)
    assert w <= 3 * ws + 1e-6, (w, ws)
    print(f"Ng={Ng} scheme={scheme} intlace={intlace}: worst vs oracle fused {w:.2e}, "
          f"separate {ws:.2e}")


def test_fused_1024_clustered_vs_oracle(psh, oracle):
    """The C2-shaped workload on the fused path (PCS, interlaced,
    N_g = 1024), a clustered box of about 1.3e6 particles."""
    dev, n = psh.gen_clustered_box(ng=256, L=1000.0, nbar=0.0013, f=0.8, seed=11)
    d = dev.to_numpy((n, 4))
    kw = dict(gsize=1024, assign=psh.PCS, intlace=True, bsize=1000.0, kbin=0.005,
              poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw))
    ctx.set_tiles(psh.fused_tile(1024, True))
    got = ctx.run_device(dev, n, float(n))
    again = ctx.run_device(dev, n, float(n))
    assert all(np.array_equal(got["pl"][0][l], again["pl"][0][l]) for l in range(3))
    ref = oracle.run(oracle.make_conf(**kw), [d])
    w = _check(got, ref, [0, 2, 4], ref["shot"][0])
    print(f"fused N_g=1024 PCS interlaced clustered: worst {w:.2e}")


def test_fused_kmax_below_nyquist(psh, oracle):
    """Only k_z up to the last live x tile is stored by the fused deposit."""
    L, n = 1000.0, 300000
    d = _uniform(n, [L] * 3, 77, weights=False)
    kw = dict(gsize=512, assign=psh.TSC, intlace=True, bsize=L, kbin=0.01, kmax=0.4,
              poles=[0, 2, 4])
    _, got = _run(psh, kw, d, psh.fused_tile(512, True))
    ref = oracle.run(oracle.make_conf(**kw), [d])
    _check(got, ref, [0, 2, 4], ref["shot"][0])


def test_fused_batch_graph(psh, oracle):
    """A batch of mocks of different sizes through the captured graph with
    the fused tiles, each mock against its own oracle run."""
    sizes = [30000, 90000, 55000]
    cats = [_uniform(s, [600.0] * 3, 880 + i, weights=True) for i, s in enumerate(sizes)]
    kw = dict(gsize=256, assign=psh.TSC, intlace=True, bsize=600.0, kbin=0.01, poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw), nmock_max=3)
    ctx.set_tiles(psh.fused_tile(256, True))
    off = np.concatenate([[0], np.cumsum(sizes)])
    dev = psh.DeviceArray.from_numpy(np.concatenate(cats))
    ws = [float(c[:, 3].sum()) for c in cats]
    first = ctx.run_batch(dev, off, ws, graph=True)
    got = ctx.run_batch(dev, off, ws, graph=True)
    for m, c in enumerate(cats):
        assert all(np.array_equal(got[m]["pl"][0][l], first[m]["pl"][0][l]) for l in range(3))
        ref = oracle.run(oracle.make_conf(**kw), [c])
        _check(got[m], ref, [0, 2, 4], ref["shot"][0])
