"""psh_run_device_static, the allocation-free variant of psh_ctx_run_device
(every intermediate buffer in the library's file-scope device scratch), run
through the C-ABI: two calls in a row, the second with a smaller catalogue,
each against psh_ctx_run_device on the same input (counts bit-exact, P_l
within 1e-9 of the bin scale: the same kernels, only the order of the FP64
bin sums may differ).  While it runs every ordinary device allocation is made
to fail, so a call that allocated would error out."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _uniform(n, L, seed):
    rng = np.random.default_rng(seed)
    d = np.empty((n, 4))
    d[:, :3] = rng.uniform(0.0, L, (n, 3))
    d[:, 3] = rng.uniform(0.5, 1.5, n)
    return d


def _same(got, want, shot):
    assert got["nbin"] == want["nbin"]
    assert np.array_equal(got["cnt"], want["cnt"])
    assert np.array_equal(got["kedge"], want["kedge"])
    scale = np.abs(want["pl"][0][0]) + shot
    for li in range(want["pl"][0].shape[0]):
        assert np.all(np.abs(got["pl"][0][li] - want["pl"][0][li]) <= 1e-9 * scale), li


def test_static_run_matches_context_run(psh):
    L = 1000.0
    kw = dict(gsize=128, assign=psh.PCS, intlace=True, bsize=L, kbin=0.01, poles=[0, 2, 4])
    conf = psh.make_conf(**kw)
    ctx = psh.Context(conf)
    nb = ctx.nbin()
    for n, seed in ((400_000, 1), (60_000, 2)):          # second call: smaller input
        d = _uniform(n, L, seed)
        dev = psh.DeviceArray.from_numpy(d)
        w = float(d[:, 3].sum())
        want = ctx.run_device(dev, n, w)
        psh.fault_inject(psh.FAULT_ALLOC, -1)             # any hipMalloc would fail
        try:
            got = psh.run_device_static(conf, dev, n, w, nb)
        finally:
            psh.fault_inject(psh.FAULT_ALLOC, 0)
        _same(got, want, L ** 3 / w)


def test_static_run_limits(psh):
    L = 500.0
    conf = psh.make_conf(gsize=64, assign=psh.CIC, bsize=L, kbin=0.02, poles=[0, 2])
    nb = psh.Context(conf).nbin()
    d = _uniform(1000, L, 3)
    dev = psh.DeviceArray.from_numpy(d)
    with pytest.raises(psh.PshError) as e:                # output too small
        psh.run_device_static(conf, dev, len(d), float(d[:, 3].sum()), nb - 1)
    assert e.value.code == psh.ERR_ARG
    big = psh.make_conf(gsize=1024, assign=psh.CIC, bsize=L, kbin=0.02, poles=[0, 2])
    with pytest.raises(psh.PshError) as e:                # meshes beyond the scratch
        psh.run_device_static(big, dev, len(d), float(d[:, 3].sum()), 4096)
    assert e.value.code == psh.ERR_MEMORY
    assert "static device scratch" in str(e.value)
    # the scratch keeps the last call's contents; a following call is unaffected
    ref = psh.Context(conf).run_device(dev, len(d), float(d[:, 3].sum()))
    got = psh.run_device_static(conf, dev, len(d), float(d[:, 3].sum()), nb)
    _same(got, ref, L ** 3 / float(d[:, 3].sum()))
