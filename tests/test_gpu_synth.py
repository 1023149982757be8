"""GPU generators draw the host generators' streams: Philox words
bit-identical, uniform boxes bit-identical, sky samples within FP64 ulps,
clustered boxes identical cell by cell wherever the FP32 (GPU) and FP64
(host) Gaussian fields give the same Poisson count."""
import numpy as np
import pytest

from powspec_amd import synth as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,stream,draw", [(0, 0, 0), (2 ** 40 + 17, 3, 0), (7, 5, 1)])
def test_philox_words(psh, seed, stream, draw):
    n = 100000
    got = psh.gen_philox(n, seed, stream, draw)
    ref = np.stack(S.draw(seed, np.arange(n, dtype=np.uint64), stream, draw), axis=1)
    assert np.array_equal(got.astype(np.uint64), ref)


def test_uniform_box_identical(psh):
    n = 300000
    dev = psh.gen_uniform_box(n, [1000.0, 700.0, 1234.5], 21)
    got = dev.to_numpy((n, 4))
    ref = S.uniform_box(n, [1000.0, 700.0, 1234.5], 21)
    assert np.array_equal(got, ref)


def test_sky_matches_host(psh):
    n = 200000
    d, nz = psh.gen_sky(n, 31)
    got, gnz = d.to_numpy((n, 4)), nz.to_numpy((n,))
    ref, rnz = S.sky(n, 31)
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-12)
    assert np.allclose(gnz, rnz, rtol=1e-12)


def test_clustered_box_matches_host(psh):
    kw = dict(ng=32, L=400.0, nbar=0.01, f=0.8, seed=4)
    dev, n = psh.gen_clustered_box(**kw)
    got = dev.to_numpy((n, 4))
    ref = S.clustered_box(ng_gen=32, L=400.0, nbar=0.01, f=0.8, seed=4)
    assert abs(n - len(ref)) <= 1e-4 * len(ref) + 3
    assert got[:, :3].min() >= 0 and got[:, :3].max() < 400.0
    # same (x, y) draws for the particles of every cell whose count agrees
    g = {tuple(r) for r in np.round(got[:, :2], 9)}
    match = np.mean([tuple(r) in g for r in np.round(ref[:, :2], 9)])
    assert match > 0.999
    # z carries f Psi_z from FP32 vs FP64 fields
    order_g = np.lexsort((got[:, 1], got[:, 0]))
    order_r = np.lexsort((ref[:, 1], ref[:, 0]))
    if n == len(ref):
        dz = np.abs(got[order_g, 2] - ref[order_r, 2])
        dz = np.minimum(dz, 400.0 - dz)
        assert np.quantile(dz, 0.999) < 1e-3


def test_clustered_box_feeds_the_path(psh):
    """The generated device catalogue goes straight into psh_ctx_run_device."""
    dev, n = psh.gen_clustered_box(ng=64, L=600.0, nbar=0.005, seed=8)
    ctx = psh.Context(psh.make_conf(gsize=64, assign=2, intlace=True, bsize=600.0, kbin=0.02,
                                    poles=[0, 2]))
    res = ctx.run_device(dev, n, float(n))
    assert res["nbin"] > 0 and np.all(np.isfinite(res["pl"][0]))
    assert res["pl"][0][0][2] > 0   # clustered: positive monopole on large scales
