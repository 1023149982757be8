"""Grid sizes that are not powers of two (3 * 2^k and 5 * 2^k) against the
FP64 oracle, through the C-ABI.  The reference takes any GRID_SIZE
(src/load_conf.c:1018-1030 checks only its range; FFTW plans every length);
these grids run the mixed-radix transforms (a radix-3 or radix-5 last pass
after the power-of-two passes) and deposit tiles no larger than the grid's
power-of-two factor.  The oracle transforms such a length by one
decimation-in-time step by the odd factor over radix-2 transforms (pinned
against numpy in test_oracle_math).

Bar as in test_gpu_parity: counts bit-exact, |dP_l| <= 1e-4 S_b per bin."""
import numpy as np
import pytest

from tests.test_gpu_parity import _check, _uniform

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("Ng", [48, 80, 96, 160])
@pytest.mark.parametrize("scheme,intlace", [(1, False), (1, True), (3, True)])
def test_sim_mixed_radix_grids(psh, oracle, Ng, scheme, intlace):
    L, n = 600.0, 40000
    d = _uniform(n, [L] * 3, 300 + Ng + scheme, weights=True)
    kw = dict(gsize=Ng, assign=scheme, intlace=intlace, bsize=L, kbin=0.01,
              poles=[0, 1, 2, 3, 4], los=[0.6, 0.0, 0.8])
    got = psh.Context(psh.make_conf(**kw)).run_device(psh.DeviceArray.from_numpy(d), n,
                                                      d[:, 3].sum())
    ref = oracle.run(oracle.make_conf(**kw), [d])
    w = _check(got, ref, [0, 1, 2, 3, 4], ref["shot"][0])
    print(f"Ng={Ng} scheme={scheme} intlace={intlace}: {ref['nbin']} bins, worst {w:.2e}",
          flush=True)


@pytest.mark.parametrize("Ng,scheme,intlace,kbin", [(768, 3, True, 0.005), (1536, 1, False, 0.01)])
def test_sim_large_mixed_radix_grids(psh, oracle, Ng, scheme, intlace, kbin):
    """Full-size grids of 3 * 2^8 and 3 * 2^9 cells (a radix-3 last pass on
    N_g = 768 and 1536 with the streaming x, y and z kernels) against the
    oracle."""
    L, n = 1000.0, 2_000_000
    d = _uniform(n, [L] * 3, Ng, weights=False)
    kw = dict(gsize=Ng, assign=scheme, intlace=intlace, bsize=L, kbin=kbin, poles=[0, 2, 4])
    got = psh.Context(psh.make_conf(**kw)).run_device(psh.DeviceArray.from_numpy(d), n,
                                                      float(n))
    print(f"Ng={Ng}: GPU done, {got['nbin']} bins", flush=True)
    ref = oracle.run(oracle.make_conf(**kw), [d])
    w = _check(got, ref, [0, 2, 4], ref["shot"][0])
    print(f"Ng={Ng} scheme={scheme} intlace={intlace}: worst {w:.2e}", flush=True)


def test_sim_mixed_radix_rect_box(psh, oracle):
    """A non-cubic box on a 3 * 2^5 grid (per-axis k spacings and alias
    factors)."""
    Ng, L, n = 96, [500.0, 700.0, 900.0], 50000
    d = _uniform(n, L, 7, weights=False)
    kw = dict(gsize=Ng, assign=2, intlace=True, bsize=L, kbin=0.008, poles=[0, 2, 4])
    got = psh.Context(psh.make_conf(**kw)).run_device(psh.DeviceArray.from_numpy(d), n,
                                                      float(n))
    ref = oracle.run(oracle.make_conf(**kw), [d])
    _check(got, ref, [0, 2, 4], ref["shot"][0])


def test_batched_mixed_radix(psh, oracle):
    """Three mocks of different sizes in one graph replay on a 5 * 2^4 grid."""
    Ng, L = 80, 400.0
    cats = [_uniform(n, [L] * 3, 900 + i, weights=True) for i, n in
            enumerate([5000, 40000, 17000])]
    kw = dict(gsize=Ng, assign=psh.TSC, intlace=True, bsize=L, kbin=0.02, poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**kw), nmock_max=3)
    off = np.concatenate([[0], np.cumsum([len(c) for c in cats])])
    dev = psh.DeviceArray.from_numpy(np.concatenate(cats))
    ws = [float(c[:, 3].sum()) for c in cats]
    ctx.run_batch(dev, off, ws, graph=True)
    got = ctx.run_batch(dev, off, ws, graph=True)
    for m, c in enumerate(cats):
        ref = oracle.run(oracle.make_conf(**kw), [c])
        _check(got[m], ref, [0, 2, 4], ref["shot"][0])


def test_survey_mixed_radix(psh, oracle):
    """Survey estimator (Y_lm meshes, l = 0, 2, 4) on a 3 * 2^5 grid."""
    rng = np.random.default_rng(11)
    lo, hi = np.array([100.0, -200.0, 300.0]), np.array([500.0, 150.0, 700.0])
    r = np.empty((150000, 4))
    r[:, :3] = lo + (hi - lo) * rng.uniform(0, 1, (150000, 3))
    r[:, 3] = 1.0
    d = np.empty((25000, 4))
    d[:, :3] = lo + (hi - lo) * rng.uniform(0, 1, (25000, 3)) ** 1.3
    d[:, 3] = 1.0
    alpha = len(d) / len(r)
    sums = {"wdata": float(len(d)), "wrand": float(len(r)), "alpha": alpha,
            "shot": len(d) + alpha * alpha * len(r), "norm": alpha * len(r) * 3e-4}
    kw = dict(gsize=96, assign=psh.CIC, intlace=True, issim=False, kbin=0.01, poles=[0, 2, 4])
    got = psh.Context(psh.make_conf(**kw)).run_host([d], [r], [sums])
    ref = oracle.run(oracle.make_conf(**kw), [d], [r], [sums])
    assert ref["err"] == 0
    _check(got, ref, [0, 2, 4], sums["shot"], sums["norm"])


@pytest.mark.parametrize("Ng", [12, 100, 112, 3072])
def test_unsupported_grid_sizes_rejected(psh, Ng):
    with pytest.raises(psh.PshError) as e:
        psh.Context(psh.make_conf(gsize=Ng, assign=1, bsize=100.0, kbin=0.01))
    assert "GRID_SIZE" in str(e.value)
