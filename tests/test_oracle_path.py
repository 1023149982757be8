"""Pins of the oracle's path functions (CPU only): assignment windows, box
definition, mode counts, interlacing, binning, the survey multipole fields and
the composed pipeline, each against an independent numpy brute force."""
import math

import numpy as np
import pytest
import scipy.special as sps


# ---- textbook windows (README.md:55-70 of the reference) ---------------------

def _win(scheme, d):
    a = np.abs(d)
    if scheme == 0:
        return ((d >= -0.5) & (d < 0.5)).astype(float)
    if scheme == 1:
        return np.where(a < 1, 1 - a, 0.0)
    if scheme == 2:
        return np.where(a < 0.5, 0.75 - d * d,
                        np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))
    return np.where(a < 1, (4 - 6 * d * d + 3 * a ** 3) / 6,
                    np.where(a < 2, (2 - a) ** 3 / 6, 0.0))


def brute_mesh(scheme, data, Ng, bmin, L):
    """sum_p w_p prod_axes W(d) with periodic grid distances."""
    x = (data[:, :3] - np.asarray(bmin)) * Ng / np.asarray(L)
    g = np.arange(Ng)
    W = []
    for a in range(3):
        d = x[:, a, None] - g[None, :]
        d = (d + Ng / 2) % Ng - Ng / 2
        W.append(_win(scheme, d))
    return np.einsum("p,pi,pj,pk->ijk", data[:, 3], W[0], W[1], W[2])


def _particles(n, L, seed, w=None):
    rng = np.random.default_rng(seed)
    d = np.empty((n, 4))
    d[:, :3] = rng.uniform(0, 1, (n, 3)) * np.asarray(L)
    d[:, 3] = 1.0 if w is None else rng.uniform(0.5, 2.0, n)
    return d


@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_assign_matches_textbook_windows(oracle, scheme):
    Ng, L = 8, [100.0, 120.0, 80.0]
    d = _particles(300, L, 10 + scheme, w=True)
    got = oracle.assign(scheme, d, Ng, [0, 0, 0], L)
    ref = brute_mesh(scheme, d, Ng, [0, 0, 0], L)
    assert np.allclose(got, ref, atol=1e-12)
    assert math.isclose(got.sum(), d[:, 3].sum(), rel_tol=1e-12)   # mass


@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_assign_translation_by_cells(oracle, scheme):
    Ng, L = 16, 160.0
    d = _particles(200, [L] * 3, 20 + scheme)
    d[:, :3] *= 0.5
    a = oracle.assign(scheme, d, Ng, [0, 0, 0], [L] * 3)
    d2 = d.copy()
    d2[:, :3] += np.array([3, 1, 2]) * L / Ng
    b = oracle.assign(scheme, d2, Ng, [0, 0, 0], [L] * 3)
    assert np.allclose(np.roll(a, (3, 1, 2), axis=(0, 1, 2)), b, atol=1e-9)


def test_assign_ngp_node_and_wraps(oracle):
    Ng, L = 8, 8.0
    d = np.array([[3.0, 4.0, 7.6, 2.5]])     # z rounds up past the edge -> 0
    m = oracle.assign(0, d, Ng, [0, 0, 0], [L] * 3)
    assert m[3, 4, 0] == 2.5 and m.sum() == 2.5
    d = np.array([[7.9, 0.05, 3.2, 1.0]])    # PCS touches x=6,7,0,1 and y=7,0,1,2
    m = oracle.assign(3, d, Ng, [0, 0, 0], [L] * 3)
    nzx = sorted(set(np.argwhere(m > 0)[:, 0]))
    nzy = sorted(set(np.argwhere(m > 0)[:, 1]))
    assert nzx == [0, 1, 6, 7] and nzy == [0, 1, 2, 7]


# ---- box definition (src/genr_mesh.c:509-578) -------------------------------

def test_def_box(oracle):
    err, size, bmin = oracle.def_box([1, 2, 3], [5, 6, 7], True, [10, 10, 10], [0.02] * 3)
    assert err == 0 and list(size) == [10] * 3 and list(bmin) == [0] * 3
    err, _, _ = oracle.def_box([-1, 2, 3], [5, 6, 7], True, [10] * 3, [0.02] * 3)
    assert err == -11
    err, _, _ = oracle.def_box([1, 2, 3], [5, 6, 10], True, [10] * 3, [0.02] * 3)
    assert err == -11
    # survey, automatic: ceil(extent*(1+pad)/10)*10, centred
    err, size, bmin = oracle.def_box([-100, 0, 50], [300, 7, 1049], False, None, [0.02] * 3)
    assert err == 0
    assert list(size) == [410.0, 10.0, 1020.0]
    assert list(bmin) == [(300 - 100 - 410) / 2, (7 - 10) / 2, (1049 + 50 - 1020) / 2]
    # survey with BOX_SIZE
    err, size, bmin = oracle.def_box([0, 0, 0], [10, 20, 30], False, [50, 50, 50], [0] * 3)
    assert err == 0 and list(bmin) == [-20, -15, -10]
    err, _, _ = oracle.def_box([0, 0, 0], [10, 20, 60], False, [50, 50, 50], [0] * 3)
    assert err == -11


# ---- mode counts and mean k (src/multipole.c:111-257) -------------------------

def _full_grid_kmod(Ng, L):
    s = np.where(np.arange(Ng) <= Ng // 2, np.arange(Ng), np.arange(Ng) - Ng).astype(float)
    v = 2 * np.pi / np.asarray(L, float)
    ki, kj, kk = (s * v[0])[:, None, None], (s * v[1])[None, :, None], (s * v[2])[None, None, :]
    # same association as the reference: ((ki^2 + kj^2) + kk^2)
    return (ki * ki + kj * kj) + kk * kk, ki, kj, kk


@pytest.mark.parametrize("Ng,L,dk", [(16, [1000.0] * 3, 0.01), (12, [300.0, 420.0, 350.0], 0.02),
                                     (32, [1000.0] * 3, 0.005)])
def test_precomp_counts_full_grid_brute_force(oracle, Ng, L, dk):
    nb, kedge, _ = oracle.kbins(Ng, L, False, 0.0, -1.0, dk)
    alias, cnt, km = oracle.precomp(Ng, L, False, 1, False, kedge, dk)
    k2, *_ = _full_grid_kmod(Ng, L)
    kmod = np.sqrt(k2)
    inr = (kmod >= kedge[0]) & (kmod < kedge[-1])
    b = ((kmod - kedge[0]) / dk).astype(np.int64)
    inr &= (b >= 0) & (b < nb)
    ref_cnt = np.bincount(b[inr], minlength=nb)
    assert np.array_equal(cnt.astype(np.int64), ref_cnt)
    # every mode of the full grid is either counted or out of range
    assert int(cnt.sum()) + int((~inr).sum()) == Ng ** 3
    ref_km = np.bincount(b[inr], weights=kmod[inr], minlength=nb)
    nzb = ref_cnt > 0
    assert np.allclose(km[nzb], ref_km[nzb] / ref_cnt[nzb], rtol=1e-13)
    # DC mode lands in bin 0 when kmin = 0 (quirk kept)
    assert alias[0, 0, 0] != 0


# ---- interlacing (src/multipole.c:461-487) --------------------------------------

def test_interlace_half_cell_shift_band_limited(oracle):
    # B is A sampled half a cell lower; for a band-limited field the combined
    # spectrum equals A's spectrum.
    Ng = 16
    g = np.arange(Ng, dtype=float)
    i, j, k = np.meshgrid(g, g, g, indexing="ij")
    waves = [((2, -3, 1), 1.0), ((5, 4, 6), 0.5), ((-1, 0, 3), 0.25)]
    A = sum(a * np.cos(2 * np.pi * (q[0] * i + q[1] * j + q[2] * k) / Ng) for q, a in waves)
    B = sum(a * np.cos(2 * np.pi * (q[0] * (i - .5) + q[1] * (j - .5) + q[2] * (k - .5)) / Ng)
            for q, a in waves)
    FA = oracle.fft_r2c(A)
    comb = oracle.interlace(FA, oracle.fft_r2c(B))
    assert np.max(np.abs(comb - FA)) < 1e-10 * np.max(np.abs(FA))


# ---- binning -------------------------------------------------------------------

def test_count_sim_constant_field_full_grid(oracle):
    Ng, L, dk = 16, [1000.0] * 3, 0.02
    poles = [0, 2, 4]
    nb, kedge, _ = oracle.kbins(Ng, L, False, 0.0, -1.0, dk)
    alias, cnt, _ = oracle.precomp(Ng, L, False, 0, False, kedge, dk)   # NGP: alias 1
    F = np.ones((Ng, Ng, Ng // 2 + 1), complex)
    pl, lcnt = oracle.count_sim(Ng, L, False, kedge, dk, [0, 0, 1], poles, alias, F, F)
    k2, _, _, kk = _full_grid_kmod(Ng, L)
    kmod = np.sqrt(k2)
    b = ((kmod - kedge[0]) / dk).astype(np.int64)
    ok = (kmod >= kedge[0]) & (kmod < kedge[-1]) & (k2 > 0)
    mu = np.divide(np.broadcast_to(kk, kmod.shape), kmod, out=np.zeros_like(kmod), where=k2 > 0)
    for li, ell in enumerate(poles):
        ref = np.bincount(b[ok], weights=sps.eval_legendre(ell, mu[ok]), minlength=nb)
        assert np.allclose(pl[li], ref, atol=1e-9), ell
        assert np.allclose(lcnt[li], ref, atol=1e-9), ell


def test_count_survey_random_field_full_grid(oracle):
    Ng, L, dk = 16, [800.0, 900.0, 1000.0], 0.02
    nb, kedge, _ = oracle.kbins(Ng, L, False, 0.0, -1.0, dk)
    alias, _, _ = oracle.precomp(Ng, L, False, 0, False, kedge, dk)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((Ng, Ng, Ng))
    F = np.fft.rfftn(x)
    pl = oracle.count_survey(Ng, L, False, kedge, dk, alias, F, F)
    Ff = np.fft.fftn(x)
    k2, *_ = _full_grid_kmod(Ng, L)
    kmod = np.sqrt(k2)
    b = ((kmod - kedge[0]) / dk).astype(np.int64)
    ok = (kmod >= kedge[0]) & (kmod < kedge[-1])
    ref = np.bincount(b[ok], weights=np.abs(Ff[ok]) ** 2, minlength=nb)
    assert np.allclose(pl, ref, rtol=1e-12)


# ---- survey multipole fields (src/mp_template.c:48-149) -----------------------

def _ysum(ell, v):
    """sum_m of the real-form Y_lm at unit vectors v (scipy based)."""
    th = np.arccos(np.clip(v[..., 2], -1, 1))
    ph = np.arctan2(v[..., 1], v[..., 0])
    s = sps.sph_harm_y(ell, 0, th, ph).real
    for m in range(1, ell + 1):
        Y = sps.sph_harm_y(ell, m, th, ph)
        s = s + math.sqrt(2) * (-1) ** m * (Y.real + Y.imag)
    return s


@pytest.mark.parametrize("ell", [2, 4])
def test_dens_kl_brute_force_dft(oracle, ell):
    """F_l(k) = sum_m Y_lm(k) FFT[F Y_lm(r)] equals, by the addition theorem,
    (2l+1)/(4 pi) sum_r F(r) L_l(k.r) e^{-ik.r}; rows/modes the reference treats
    specially use sum_m Y_lm of the other direction."""
    Ng = 8
    L = np.array([500.0, 600.0, 700.0])
    bmin = np.array([-300.0, 150.0, 400.0])
    smin_g = bmin * Ng / L
    rng = np.random.default_rng(ell)
    Fr = rng.standard_normal((Ng, Ng, Ng))
    nb, kedge, _ = oracle.kbins(Ng, L, False, 0.0, -1.0, 0.01)
    alias, _, _ = oracle.precomp(Ng, L, False, 1, False, kedge, 0.01)
    got = oracle.dens_kl(ell, Fr, L, smin_g, alias)

    g = np.arange(Ng)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    r = np.stack([(I + smin_g[0]) * L[0], (J + smin_g[1]) * L[1], (K + smin_g[2]) * L[2]], -1)
    rhat = r / np.linalg.norm(r, axis=-1)[..., None]
    first_row = (I == 0) & (J == 0)
    s = np.where(g <= Ng // 2, g, g - Ng)
    ref = np.zeros((Ng, Ng, Ng // 2 + 1), complex)
    for a in range(Ng):
        for b in range(Ng):
            for c in range(Ng // 2 + 1):
                if alias[a, b, c] == 0:
                    continue
                kv = np.array([s[a] / L[0], s[b] / L[1], c / L[2]])
                ph = np.exp(-2j * np.pi * (a * I + b * J + c * K) / Ng)
                if kv[0] == 0 and kv[1] == 0:
                    # unweighted in k; the first row is unweighted in r too
                    wr = np.where(first_row, 2 * ell + 1, _ysum(ell, rhat))
                else:
                    khat = kv / np.linalg.norm(kv)
                    leg = (2 * ell + 1) / (4 * np.pi) * sps.eval_legendre(ell, rhat @ khat)
                    wr = np.where(first_row, _ysum(ell, khat[None, :])[0], leg)
                ref[a, b, c] = np.sum(Fr * wr * ph)
    assert np.allclose(got, ref, atol=1e-9 * np.abs(ref).max())


# ---- the composed pipeline -------------------------------------------------------

def _numpy_sim_pipeline(data, Ng, L, dk, poles):
    """NGP, no interlacing (alias == 1): numpy mesh -> rfftn -> full-grid bins."""
    mesh = brute_mesh(0, data, Ng, [0, 0, 0], [L] * 3)
    F = np.fft.fftn(mesh)
    k2, _, _, kk = _full_grid_kmod(Ng, [L] * 3)
    kmod = np.sqrt(k2)
    nb = int(round(math.pi * Ng / L / dk))
    if dk * nb > math.pi * Ng / L:
        nb -= 1
    kedge = dk * np.arange(nb + 1)
    b = ((kmod - kedge[0]) / dk).astype(np.int64)
    inr = (kmod >= kedge[0]) & (kmod < kedge[-1])
    cnt = np.bincount(b[inr], minlength=nb)
    ok = inr & (k2 > 0)
    mu = np.divide(np.broadcast_to(kk, kmod.shape), kmod, out=np.zeros_like(kmod), where=k2 > 0)
    W = data[:, 3].sum()
    V = L ** 3
    shot, norm = V / W, W * W / V
    out = []
    for ell in poles:
        leg = sps.eval_legendre(ell, mu[ok])
        s = np.bincount(b[ok], weights=np.abs(F[ok]) ** 2 * leg, minlength=nb)
        lc = np.bincount(b[ok], weights=leg, minlength=nb)
        out.append((s / (norm * cnt) - shot * lc / cnt) * (2 * ell + 1))
    return cnt, np.array(out)


def test_run_sim_matches_numpy_composition(oracle):
    Ng, L, dk = 16, 1000.0, 0.01
    d = _particles(400, [L] * 3, 77, w=True)
    conf = oracle.make_conf(gsize=Ng, assign=0, intlace=False, bsize=L, kbin=dk)
    r = oracle.run(conf, [d])
    assert r["err"] == 0
    cnt, P = _numpy_sim_pipeline(d, Ng, L, dk, [0, 2, 4])
    assert np.array_equal(r["cnt"].astype(np.int64), cnt)
    assert np.allclose(r["pl"][0], P, rtol=1e-10, atol=1e-6 * np.abs(P).max())


def test_run_uniform_box_is_shot_noise(oracle):
    # Poisson sampling: P_0 ~ 0 after shot-noise subtraction, P_2, P_4 ~ 0.
    Ng, L, n = 32, 1000.0, 200000
    d = _particles(n, [L] * 3, 5)
    conf = oracle.make_conf(gsize=Ng, assign=3, intlace=True, bsize=L, kbin=0.01)
    r = oracle.run(conf, [d])
    shot = r["shot"][0]
    assert shot == pytest.approx(L ** 3 / n)
    P = r["pl"][0]
    good = r["cnt"] > 200
    for li in range(3):
        rel = np.abs(P[li][good]) / shot
        assert np.median(rel) < 0.2, li


def test_run_errors(oracle):
    d = _particles(10, [100.0] * 3, 1)
    d[0, 0] = 100.0   # not < BOX_SIZE
    conf = oracle.make_conf(gsize=8, bsize=100.0, kbin=0.05)
    assert oracle.run(conf, [d])["err"] == -13
    d = _particles(10, [100.0] * 3, 1)
    conf = oracle.make_conf(gsize=4, bsize=100.0, kbin=1.0)   # no k bins
    assert oracle.run(conf, [d])["err"] == -14
