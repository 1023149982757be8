"""Host generators (powspec_amd/synth.py): the Philox4x32-10 known-answer
vectors of Salmon et al. (2011), uniformity, the Gaussian field's power
against the input P(k) within cosmic variance, Poisson sampling moments and
the survey sky distributions."""
import numpy as np
import pytest
from scipy import stats

from powspec_amd import synth as S

KAT = [([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
       ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
       ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1])]


@pytest.mark.parametrize("ctr,key,exp", KAT)
def test_philox_known_answers(ctr, key, exp):
    got = S.philox4x32_10(*[np.array([c], np.uint64) for c in ctr], *key)
    assert [int(g[0]) for g in got] == exp


def test_u53_range_and_uniform_box():
    d = S.uniform_box(200000, 1000.0, 5)
    assert d[:, :3].min() >= 0 and d[:, :3].max() < 1000.0
    assert np.all(d[:, :3] * 2048 / 1000.0 < 2048)
    for a in range(3):
        assert stats.kstest(d[:, a] / 1000.0, "uniform").pvalue > 1e-4
    assert np.all(d[:, 3] == 1.0)
    # independent seeds and axes
    e = S.uniform_box(200000, 1000.0, 6)
    assert abs(np.corrcoef(d[:, 0], e[:, 0])[0, 1]) < 0.01
    assert abs(np.corrcoef(d[:, 0], d[:, 1])[0, 1]) < 0.01


def test_normals_and_poisson():
    z = S.normals(9, np.arange(400000, dtype=np.uint64))
    assert stats.kstest(z, "norm").pvalue > 1e-4
    lam = np.full(300000, 3.7)
    u = S.u53(*S.draw(9, np.arange(300000, dtype=np.uint64), S.POISSON, 0)[:2])
    k = S.poisson_inv(lam, u)
    assert abs(k.mean() - 3.7) < 0.02 and abs(k.var() - 3.7) < 0.05
    assert np.array_equal(S.poisson_inv(np.array([0.0, 2.0]), np.array([0.5, 0.0])), [0, 0])


def test_gaussian_field_power():
    """Shell-averaged |delta_k|^2 V / N^2 matches the input P(k): the mean
    ratio over the well-sampled shells is 1 within cosmic variance."""
    ng, L = 64, 800.0
    delta, _ = S.gaussian_field(ng, L, 11)
    dk = np.fft.rfftn(delta.astype(np.float64))
    kf = 2 * np.pi / L
    kx = np.fft.fftfreq(ng, d=1.0 / ng) * kf
    kz = np.arange(ng // 2 + 1) * kf
    k = np.sqrt(kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2)
    pm = np.abs(dk) ** 2 * L ** 3 / ng ** 6
    edges = np.arange(2, 20) * kf * 1.5
    ratios, nmodes = [], []
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = (k >= lo) & (k < hi)
        ratios.append(pm[sel].mean() / S.power_law_cut(k[sel]).mean())
        nmodes.append(sel.sum())
    ratios = np.array(ratios)
    # each shell: relative error ~ 1/sqrt(modes); all shells together
    assert np.all(np.abs(ratios - 1) < 5 / np.sqrt(np.array(nmodes)) + 0.02)
    assert abs(np.average(ratios, weights=nmodes) - 1) < 0.02


def test_clustered_box_counts_and_shift():
    d, counts = S.clustered_box(ng_gen=32, L=400.0, nbar=0.01, seed=4, return_counts=True)
    assert len(d) == counts.sum()
    exp = 0.01 * 400.0 ** 3
    assert abs(len(d) - exp) < 6 * np.sqrt(exp) + 0.05 * exp   # lognormal mean one
    assert d[:, :3].min() >= 0 and d[:, :3].max() < 400.0
    # clustering: cell counts are over-dispersed relative to Poisson
    assert counts.var() > 1.5 * counts.mean()
    # without RSD the particles stay in their cells
    e, c2 = S.clustered_box(ng_gen=32, L=400.0, nbar=0.01, seed=4, f=0.0, return_counts=True)
    assert np.array_equal(counts, c2)
    cells = np.repeat(np.arange(32 ** 3), c2)
    ijk = np.floor(e[:, :3] / (400.0 / 32)).astype(int)
    assert np.array_equal(ijk[:, 0] * 1024 + ijk[:, 1] * 32 + ijk[:, 2], cells)
    assert np.array_equal(d[:, :2], e[:, :2])   # RSD moves z only


def test_sky_distributions():
    d, nz = S.sky(200000, 12)
    assert d[:, 0].min() >= 0 and d[:, 0].max() < 60
    s = np.sin(np.radians(d[:, 1]))
    assert stats.kstest((s - np.sin(np.radians(-30))) / (2 * np.sin(np.radians(30))),
                        "uniform").pvalue > 1e-4
    tn = stats.truncnorm((0.4 - 0.55) / 0.08, (0.7 - 0.55) / 0.08, loc=0.55, scale=0.08)
    assert stats.kstest(d[:, 2], tn.cdf).pvalue > 1e-4
    assert np.allclose(d[:, 3], 1 / (1 + nz * 1e4))
