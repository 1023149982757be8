"""Forced failures on the GPU, injected from the host (psh_fault_inject): every
entry point returns the documented status code instead of exiting (the
reference exit()s inside parallel regions on allocation failure,
src/multipole.c:121-126), and a context whose allocation failed keeps working
correctly once memory is available again: no stale buffer capacity and no
half-built box survives the failure.  No kernel runs with an invalid pointer:
the failures happen on the host before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _uniform(n, L, seed):
    rng = np.random.default_rng(seed)
    d = np.empty((n, 4))
    d[:, :3] = rng.uniform(0.0, L, (n, 3))
    d[:, 3] = 1.0
    return d


@pytest.fixture
def faults(psh):
    yield psh
    psh.fault_inject(psh.FAULT_ALLOC, 0)
    psh.fault_inject(psh.FAULT_GRAPH, 0)


def _close(a, b, rel=1e-9):
    assert np.array_equal(a["cnt"], b["cnt"])
    scale = np.abs(b["pl"][0][0]) + 1e-30
    for li in range(a["pl"][0].shape[0]):
        assert np.all(np.abs(a["pl"][0][li] - b["pl"][0][li]) <= rel * scale)


def test_create_fails_at_every_allocation(faults):
    P = faults
    conf = P.make_conf(gsize=32, assign=P.TSC, intlace=True, bsize=500.0, kbin=0.02)
    k, hit = 1, 0
    while True:
        P.fault_inject(P.FAULT_ALLOC, k)
        try:
            ctx = P.Context(conf)
        except P.PshError as e:
            assert e.code == P.ERR_MEMORY, e
            assert "injected" in str(e)
            hit += 1
            k += 1
            continue
        P.fault_inject(P.FAULT_ALLOC, 0)
        ctx.close()
        break
    assert hit >= 5   # meshes, tables, bins: every one of them was failed once


def test_run_fails_then_recovers(faults):
    """Each allocation of a first run fails once in turn; the same context
    then runs correctly (compared with a fresh context)."""
    P = faults
    L, n = 500.0, 20000
    d = _uniform(n, L, 3)
    dev = P.DeviceArray.from_numpy(d)
    conf = P.make_conf(gsize=32, assign=P.PCS, intlace=True, bsize=L, kbin=0.02)
    ref = P.Context(conf).run_device(dev, n, float(n))
    failures = 0
    for k in range(1, 40):
        ctx = P.Context(conf)
        P.fault_inject(P.FAULT_ALLOC, k)
        try:
            ctx.run_device(dev, n, float(n))
            ok = True
        except P.PshError as e:
            assert e.code == P.ERR_MEMORY, e
            ok = False
            failures += 1
        P.fault_inject(P.FAULT_ALLOC, 0)
        _close(ctx.run_device(dev, n, float(n)), ref)
        ctx.close()
        if ok:
            break
    assert failures >= 3


def test_batch_growth_failure_then_smaller_batch(faults):
    """A batch that needs larger scratch fails to allocate it; a smaller batch
    on the same context then gives the same spectra as before, and the large
    batch succeeds once memory is back (the advisor's stale-capacity case)."""
    P = faults
    L = 400.0
    sizes = [3000, 5000, 40000, 60000]
    parts = [_uniform(s, L, 10 + i) for i, s in enumerate(sizes)]
    conf = P.make_conf(gsize=32, assign=P.CIC, intlace=True, bsize=L, kbin=0.02)
    ctx = P.Context(conf, nmock_max=4)
    small = np.concatenate(parts[:2])
    ds = P.DeviceArray.from_numpy(small)
    off_s = [0, sizes[0], sizes[0] + sizes[1]]
    first = ctx.run_batch(ds, off_s, sizes[:2])
    big = np.concatenate(parts)
    db = P.DeviceArray.from_numpy(big)
    off_b = np.concatenate([[0], np.cumsum(sizes)])
    P.fault_inject(P.FAULT_ALLOC, -1)          # every allocation fails
    with pytest.raises(P.PshError) as e:
        ctx.run_batch(db, off_b, sizes)
    assert e.value.code == P.ERR_MEMORY
    P.fault_inject(P.FAULT_ALLOC, 0)
    again = ctx.run_batch(ds, off_s, sizes[:2])
    for a, b in zip(again, first):
        _close(a, b)
    full = ctx.run_batch(db, off_b, sizes)
    fresh = P.Context(conf, nmock_max=4).run_batch(db, off_b, sizes)
    for a, b in zip(full, fresh):
        _close(a, b)


def test_set_box_failure_leaves_no_half_box(faults):
    P = faults
    L, n = 300.0, 10000
    d = _uniform(n, L, 5)
    conf = P.make_conf(gsize=32, assign=P.CIC, bsize=L, kbin=0.02)
    ctx = P.Context(conf)
    dev = P.DeviceArray.from_numpy(d)
    P.fault_inject(P.FAULT_ALLOC, 1)
    with pytest.raises(P.PshError) as e:
        ctx.set_box([2 * L] * 3)
    assert e.value.code == P.ERR_MEMORY
    assert ctx.nbin() == 0                    # no box until one is complete
    P.fault_inject(P.FAULT_ALLOC, 0)
    ctx.set_box([L] * 3)
    _close(ctx.run_device(dev, n, float(n)), P.Context(conf).run_device(dev, n, float(n)))


def test_graph_instantiation_failure(faults):
    P = faults
    L = 400.0
    sizes = [4000, 6000]
    d = np.concatenate([_uniform(s, L, 20 + i) for i, s in enumerate(sizes)])
    dev = P.DeviceArray.from_numpy(d)
    off = [0, sizes[0], sum(sizes)]
    conf = P.make_conf(gsize=32, assign=P.TSC, intlace=True, bsize=L, kbin=0.02)
    ctx = P.Context(conf, nmock_max=2)
    P.fault_inject(P.FAULT_GRAPH, 1)
    with pytest.raises(P.PshError) as e:
        ctx.run_batch(dev, off, sizes, graph=True)
    assert e.value.code == P.ERR_DEVICE
    P.fault_inject(P.FAULT_GRAPH, 0)
    eager = ctx.run_batch(dev, off, sizes, graph=False)
    graph = ctx.run_batch(dev, off, sizes, graph=True)
    for a, b in zip(graph, eager):
        _close(a, b)
