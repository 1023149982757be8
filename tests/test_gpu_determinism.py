"""Run-to-run reproducibility on the GPU: the meshes are integer sums
(fixed-point deposit), the FFTs have no data-dependent order, the bin sums
of the fused x pass are 64-bit integer sums per tile folded in tile order,
and the mode geometry is summed in a fixed order -- so the same input gives
the same bits, run after run and context after context (boxes, surveys and
replayed batch graphs)."""
import numpy as np
import pytest

from test_gpu_parity import _survey_cats, _uniform

pytestmark = pytest.mark.gpu


def _same(a, b):
    for key in ("cnt", "k", "kedge", "km", "lcnt"):
        assert np.array_equal(a[key], b[key]), key
    for c in range(len(a["pl"])):
        assert np.array_equal(a["pl"][c], b["pl"][c]), c
    assert np.array_equal(a["xpl"], b["xpl"])


def test_box_bitwise_repeatable(psh):
    d = _uniform(400000, [1000.0] * 3, 91, weights=True)
    dev = psh.DeviceArray.from_numpy(d)
    kw = dict(gsize=256, assign=psh.PCS, intlace=True, bsize=1000.0, kbin=0.005,
              poles=[0, 2, 4], los=[0.0, 0.6, 0.8])
    ctx = psh.Context(psh.make_conf(**kw))
    first = ctx.run_device(dev, len(d), float(d[:, 3].sum()))
    for _ in range(3):
        _same(ctx.run_device(dev, len(d), float(d[:, 3].sum())), first)
    other = psh.Context(psh.make_conf(**kw))          # geometry computed again
    _same(other.run_device(dev, len(d), float(d[:, 3].sum())), first)


def test_clustered_1024_bitwise_repeatable(psh):
    dev, n = psh.gen_clustered_box(ng=256, L=1000.0, nbar=0.005, f=0.8, seed=7)
    kw = dict(gsize=1024, assign=psh.PCS, intlace=True, bsize=1000.0, kbin=0.005,
              poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw))
    a = ctx.run_device(dev, n, float(n))
    b = ctx.run_device(dev, n, float(n))
    _same(a, b)


def test_survey_bitwise_repeatable(psh, oracle):
    (dp, dwc, dfkp, dnz), (rp, rwc, rfkp, rnz) = _survey_cats(150000, 92)
    wd, wr, sums = oracle.survey_sums(dwc, dfkp, dnz, rwc, rfkp, rnz)
    d, r = np.column_stack([dp, wd]), np.column_stack([rp, wr])
    kw = dict(gsize=64, assign=2, intlace=True, issim=False, kbin=0.01, poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw))
    a = ctx.run_host([d], [r], [sums])
    _same(ctx.run_host([d], [r], [sums]), a)
    _same(psh.Context(psh.make_conf(**kw)).run_host([d], [r], [sums]), a)


def test_batch_graph_bitwise_repeatable(psh):
    rng = np.random.default_rng(93)
    sizes = rng.integers(20000, 60000, 6)
    allp = np.concatenate([_uniform(int(s), [1000.0] * 3, 930 + i) for i, s in enumerate(sizes)])
    off = np.concatenate([[0], np.cumsum(sizes)])
    dev = psh.DeviceArray.from_numpy(allp)
    kw = dict(gsize=128, assign=psh.TSC, intlace=True, bsize=1000.0, kbin=0.01,
              poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw), nmock_max=len(sizes))
    ws = [float(s) for s in sizes]
    first = ctx.run_batch(dev, off, ws, graph=True)
    for _ in range(2):
        for a, b in zip(ctx.run_batch(dev, off, ws, graph=True), first):
            _same(a, b)
    for a, b in zip(ctx.run_batch(dev, off, ws, graph=False), first):
        _same(a, b)
