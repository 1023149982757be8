"""The batched-mock and survey workloads at their stated shapes against the
FP64 oracle (GPU), through the C-ABI:

* batched mocks: independently seeded clustered boxes (seed 1000 + i, the
  expected particle count of mock i drawn from U[2e5, 3e6] as the benchmark
  draws it), TSC, N_g = 256, interlaced, l = 0,2,4, dk = 0.01, replayed
  through the captured batch graph; each mock compared with its own oracle
  run (src/multipole.c:1179-1278), not with the plain GPU path;
* survey: the 60x60 deg sky cap, z in [0.4, 0.7] from a Gaussian n(z),
  5e6 data and 2e7 randoms generated on the GPU and converted to comoving
  coordinates on the GPU, FKP weights, CIC, N_g = 512, interlaced,
  l = 0,2,4, dk = 0.01, against the oracle on the same converted records.

Bar as in test_gpu_parity: counts bit-exact, |dP_l| <= 1e-4 S_b per bin."""
import numpy as np
import pytest

from tests.test_gpu_parity import _check

pytestmark = pytest.mark.gpu


def _mock_sizes(nmock=256, seed=1000):
    return np.random.default_rng(seed).integers(200_000, 3_000_001, nmock)


def test_batched_mocks_graph_vs_oracle(psh, oracle):
    nm = 8
    sizes = _mock_sizes()[:nm]
    kw = dict(gsize=256, assign=psh.TSC, intlace=True, bsize=1000.0, kbin=0.01,
              poles=[0, 2, 4])
    cats, counts = [], []
    for m in range(nm):
        dev, n = psh.gen_clustered_box(ng=256, L=1000.0, f=0.8, seed=1000 + m,
                                       npart=float(sizes[m]))
        cats.append(dev.to_numpy((n, 4)))
        counts.append(n)
        dev.free()
    off = np.concatenate([[0], np.cumsum(counts)])
    ctx = psh.Context(psh.make_conf(**kw), nmock_max=nm)
    dev = psh.DeviceArray.from_numpy(np.concatenate(cats))
    ws = [float(c[:, 3].sum()) for c in cats]
    ctx.run_batch(dev, off, ws, graph=True)             # capture
    got = ctx.run_batch(dev, off, ws, graph=True)       # replay
    worst = 0.0
    for m in range(nm):
        ref = oracle.run(oracle.make_conf(**kw), [cats[m]])
        w = _check(got[m], ref, [0, 2, 4], ref["shot"][0])
        worst = max(worst, w)
        print(f"mock {m}: seed {1000 + m}, {counts[m]} particles, worst |dP_l|/S_b {w:.2e}",
              flush=True)
    print(f"batched mocks vs oracle: worst {worst:.2e}")


def test_survey_sky_cap_vs_oracle(psh, oracle):
    nd, nr = 5_000_000, 20_000_000
    kw = dict(ra=(0.0, 60.0), dec=(-30.0, 30.0), zmu=0.55, zsigma=0.08, zlo=0.4, zhi=0.7,
              nbar0=5e-3, p0=1e4)
    dd, dnz = psh.gen_sky(nd, 4001, **kw)
    rr, rnz = psh.gen_sky(nr, 4002, **kw)
    psh.cnvt_device(psh.make_cnvt(omega_m=0.31), [(dd, nd), (rr, nr)])
    d = dd.to_numpy((nd, 4))
    r = rr.to_numpy((nr, 4))
    # completeness 1; column 3 holds w_fkp (src/read_cata.c:165-183 sums)
    wd, wr, sums = oracle.survey_sums(np.ones(nd), d[:, 3], dnz.to_numpy((nd,)),
                                      np.ones(nr), r[:, 3], rnz.to_numpy((nr,)))
    conf = dict(gsize=512, assign=psh.CIC, intlace=True, issim=False, kbin=0.01,
                poles=[0, 2, 4])
    ctx = psh.Context(psh.make_conf(**conf))
    got = ctx.run_survey_device(dd, nd, rr, nr, sums["alpha"], sums["shot"], sums["norm"])
    print(f"GPU done: box {ctx.nbin()} bins", flush=True)
    ref = oracle.run(oracle.make_conf(**conf), [d], [r], [sums])
    assert ref["err"] == 0
    assert np.allclose(got["kedge"], ref["kedge"], rtol=0, atol=0)
    w = _check(got, ref, [0, 2, 4], sums["shot"], sums["norm"])
    print(f"survey sky cap vs oracle: box {ref['bsize']}, {ref['nbin']} bins, "
          f"worst |dP_l|/S_b {w:.2e}")


def test_survey_first_pole_not_zero(psh, oracle):
    """poles = [2, 4] on a survey: the reference skips the first pole when it
    is not 0 (src/multipole.c:1229-1245); P_2 stays 0, P_4 is computed."""
    rng = np.random.default_rng(71)
    lo, hi = np.array([120.0, -260.0, 310.0]), np.array([620.0, 140.0, 840.0])
    r = np.empty((200000, 4))
    r[:, :3] = lo + (hi - lo) * rng.uniform(0, 1, (200000, 3))
    r[:, 3] = 1.0
    d = np.empty((30000, 4))
    d[:, :3] = lo + (hi - lo) * rng.uniform(0, 1, (30000, 3)) ** 1.5
    d[:, 3] = 1.0
    alpha = len(d) / len(r)
    sums = {"wdata": float(len(d)), "wrand": float(len(r)), "alpha": alpha,
            "shot": len(d) + alpha * alpha * len(r), "norm": alpha * len(r) * 3e-4}
    kw = dict(gsize=32, assign=psh.CIC, intlace=True, issim=False, kbin=0.01, poles=[2, 4])
    got = psh.Context(psh.make_conf(**kw)).run_host([d], [r], [sums])
    ref = oracle.run(oracle.make_conf(**kw), [d], [r], [sums])
    assert np.all(ref["pl"][0][0] == 0) and np.all(got["pl"][0][0] == 0)
    # S_b from the monopole of the same catalogues
    full = oracle.run(oracle.make_conf(**dict(kw, poles=[0, 2, 4])), [d], [r], [sums])
    has = ref["cnt"] > 0
    scale = np.abs(full["pl"][0][0]) + sums["shot"] / sums["norm"]
    assert np.all(np.abs(full["pl"][0][2] - ref["pl"][0][1])[has] <= 1e-12 * scale[has])
    assert np.all(np.abs(got["pl"][0][1] - ref["pl"][0][1])[has] <= 1e-4 * scale[has])
