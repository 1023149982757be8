"""Independent pins of the oracle functions that the GPU parity tests alone
would not catch (CPU only): the survey estimator end to end (data - alpha
randoms, interlacing with the shifted grid, the Y_lm multipole fields and the
l = 0 / l > 0 normalisations), cross-spectrum normalisations, the survey
weight sums and their n(z) fallbacks, log bins with the reference's rescale,
the interlaced box composition and shift_cat at the box edge.

Each pin is computed here from a definition, not from the oracle's code:
textbook windows with periodic distances, numpy FFTs, the addition theorem
for the multipole fields (Hand et al. 2017), closed forms for degenerate
catalogues.  A plausible mistake in the oracle -- 2 pi for 4 pi, alpha^2 for
alpha, a dropped shift, a wrong rescale -- fails one of them."""
import math

import numpy as np
import pytest
import scipy.special as sps

from tests.test_oracle_path import _particles, _ysum, brute_mesh


# ---- helpers: the half-complex grid of the reference ----------------------------

def _sidx(Ng):
    g = np.arange(Ng)
    return np.where(g <= Ng // 2, g, g - Ng)


def _half_grid(Ng, L):
    """|k|^2 in the reference's association ((kx^2 + ky^2) + kz^2), the
    half-space multiplicity and the physical k components of the stored
    modes [Ng][Ng][Ng/2+1]."""
    s = _sidx(Ng).astype(float)
    v = 2 * np.pi / np.asarray(L, float)
    kx = (s * v[0])[:, None, None]
    ky = (s * v[1])[None, :, None]
    kz = (np.arange(Ng // 2 + 1) * v[2])[None, None, :]
    k2 = (kx * kx + ky * ky) + kz * kz
    c = np.arange(Ng // 2 + 1)
    mult = np.where((c == 0) | (c == Ng // 2), 1.0, 2.0)[None, None, :]
    shape = (Ng, Ng, Ng // 2 + 1)
    return k2, np.broadcast_to(mult, shape), [np.broadcast_to(a, shape) for a in (kx, ky, kz)]


def _lin_bins(kmod, kedge, dk):
    nb = len(kedge) - 1
    b = np.floor((kmod - kedge[0]) / dk).astype(np.int64)
    ok = (kmod >= kedge[0]) & (kmod < kedge[-1]) & (b >= 0) & (b < nb)
    return b, ok


def _interlace_half(FA, FB):
    """(F_A + e^{i pi (i' + j' + k) / N} F_B) / 2 on the stored modes: the
    half-cell shift of grid B in Fourier space."""
    Ng = FA.shape[0]
    s = _sidx(Ng)
    ph = np.pi / Ng * (s[:, None, None] + s[None, :, None] + np.arange(Ng // 2 + 1)[None, None, :])
    return 0.5 * (FA + np.exp(1j * ph) * FB)


def _field_lm(ell, Fr, L, bmin, alias):
    """F_l(k) = sum_m Y_lm(k) FFT[Fr Y_lm(r)] by the addition theorem,
    (2l+1)/(4 pi) sum_r Fr(r) L_l(k.r) e^{-ik.r}, with the reference's two
    special cases: the i = j = 0 row of Fr enters unweighted for every m
    (src/mp_template.c:75-78) and modes with k_x = k_y = 0 take no Y_lm(k)
    weight (src/mp_template.c:121-124)."""
    Ng = Fr.shape[0]
    L = np.asarray(L, float)
    sg = np.asarray(bmin, float) * Ng / L
    g = np.arange(Ng)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    r = np.stack([(I + sg[0]) * L[0], (J + sg[1]) * L[1], (K + sg[2]) * L[2]], -1)
    rhat = (r / np.linalg.norm(r, axis=-1)[..., None]).reshape(-1, 3)
    first = ((I == 0) & (J == 0)).ravel()
    ysum_r = _ysum(ell, rhat)
    fr = Fr.ravel()
    s = _sidx(Ng)
    out = np.zeros((Ng, Ng, Ng // 2 + 1), complex)
    IJK = np.stack([I.ravel(), J.ravel(), K.ravel()], -1)
    for a in range(Ng):
        for b in range(Ng):
            cs = np.arange(Ng // 2 + 1)
            ph = np.exp(-2j * np.pi * (a * IJK[:, 0, None] + b * IJK[:, 1, None]
                                       + cs[None, :] * IJK[:, 2, None]) / Ng)   # [r, c]
            for c in cs:
                if alias[a, b, c] == 0:
                    continue
                kv = np.array([s[a] / L[0], s[b] / L[1], c / L[2]])
                if kv[0] == 0 and kv[1] == 0:
                    w = np.where(first, 2 * ell + 1, ysum_r)
                else:
                    kh = kv / np.linalg.norm(kv)
                    w = (2 * ell + 1) / (4 * np.pi) * sps.eval_legendre(ell, rhat @ kh)
                    w = np.where(first, _ysum(ell, kh[None, :])[0], w)
                out[a, b, c] = np.sum(fr * w * ph[:, c])
    return out


def _survey_field(data, rand, alpha, Ng, scheme, L, bmin, intlace):
    """Data - alpha randoms on the plain grid (and on the half-cell shifted
    grid, combined in Fourier space); returns (F0 on the stored modes, the
    real field the multipole weights apply to)."""
    Fr = brute_mesh(scheme, data, Ng, bmin, L) - alpha * brute_mesh(scheme, rand, Ng, bmin, L)
    FA = np.fft.rfftn(Fr)
    if not intlace:
        return FA, Fr
    smin = np.asarray(bmin) - 0.5 * np.asarray(L) / Ng
    Fs = brute_mesh(scheme, data, Ng, smin, L) - alpha * brute_mesh(scheme, rand, Ng, smin, L)
    F0 = _interlace_half(FA, np.fft.rfftn(Fs))
    # the reference transforms the combined field back (c2r / N^3) for the
    # multipole weights (src/multipole.c:489-500)
    return F0, np.fft.irfftn(F0, s=Fr.shape, axes=(0, 1, 2))


def _survey_brute(data, rand, sums, Ng, scheme, intlace, poles, dk, ref):
    """P_l of one survey catalogue pair from the definitions: P_0 =
    (<|F0|^2> - shot) / norm, P_l = (2l+1) <Re(F0 F_l*)> / norm with F_l by
    the addition theorem (Hand et al. 2017); <.> = alias-corrected,
    multiplicity-weighted bin average."""
    L, bmin = ref["bsize"], ref["bmin"]
    k2, mult, _ = _half_grid(Ng, L)
    kedge = ref["kedge"]
    b, ok = _lin_bins(np.sqrt(k2), kedge, dk)
    nb = len(kedge) - 1
    cnt = np.bincount(b[ok], weights=mult[ok], minlength=nb)
    alias = _alias_table(Ng, L, intlace, scheme)
    F0, Fr = _survey_field(data, rand, sums["alpha"], Ng, scheme, L, bmin, intlace)
    out = np.zeros((len(poles), nb))
    w = (alias * mult)[ok]
    has = cnt > 0
    c1 = np.where(has, cnt, 1.0)
    for li, ell in enumerate(poles):
        if ell == 0:
            s = np.bincount(b[ok], weights=(np.abs(F0) ** 2)[ok] * w, minlength=nb)
            out[li] = np.where(has, (s / c1 - sums["shot"]) / sums["norm"], 0.0)
        else:
            Fl = _field_lm(ell, Fr, L, bmin, np.where(ok, 1.0, 0.0))
            s = np.bincount(b[ok], weights=(F0 * np.conj(Fl)).real[ok] * w, minlength=nb)
            # (2l+1)/(4pi) is inside F_l: the reference's 4 pi / norm
            out[li] = np.where(has, 4 * np.pi * s / (sums["norm"] * c1), 0.0)
    return cnt, out


def _alias_table(Ng, L, intlace, scheme):
    """alias_corr per axis, multiplied (src/multipole.c:46-100); the per-axis
    factor itself is pinned by its closed forms in test_oracle_math."""
    from oracle import oracle as O
    s = _sidx(Ng)
    f = np.array([O.alias_corr(intlace, scheme, math.pi * i / Ng) for i in s])
    fz = np.array([O.alias_corr(intlace, scheme, math.pi * k / Ng) for k in range(Ng // 2 + 1)])
    return f[:, None, None] * f[None, :, None] * fz[None, None, :]


def _survey_catalogue(nd, nr, seed, lo=(120.0, -260.0, 310.0), hi=(390.0, 40.0, 640.0)):
    rng = np.random.default_rng(seed)
    lo, hi = np.array(lo), np.array(hi)
    d = np.empty((nd, 4))
    r = np.empty((nr, 4))
    # data clustered towards the lower corner, randoms uniform
    d[:, :3] = lo + (hi - lo) * rng.uniform(0, 1, (nd, 3)) ** 1.7
    r[:, :3] = lo + (hi - lo) * rng.uniform(0, 1, (nr, 3))
    wc_d, wc_r = rng.uniform(0.8, 1.2, nd), rng.uniform(0.9, 1.1, nr)
    nz_d, nz_r = 2e-4 + 1e-4 * rng.uniform(0, 1, nd), 2e-4 + 1e-4 * rng.uniform(0, 1, nr)
    wf_d, wf_r = 1 / (1 + nz_d * 1e4), 1 / (1 + nz_r * 1e4)
    return d, r, (wc_d, wf_d, nz_d, wc_r, wf_r, nz_r)


# ---- survey weight sums (io/read_ascii.c:875-903, src/read_cata.c:165-183) ----

def test_survey_sums_closed_forms(oracle):
    nd, nr = 700, 9000
    a, b, c = 1.25, 0.5, 3e-4          # data: w_comp, w_fkp, n(z), constant
    a2, b2, c2 = 0.75, 0.4, 2e-4       # randoms
    one = lambda n, v: np.full(n, v)
    wd, wr, s = oracle.survey_sums(one(nd, a), one(nd, b), one(nd, c),
                                   one(nr, a2), one(nr, b2), one(nr, c2))
    assert np.all(wd == a * b) and np.all(wr == a2 * b2)       # w = w_comp w_fkp
    alpha = nd * a / (nr * a2)                                  # sum w_comp ratio
    assert s["wdata"] == pytest.approx(nd * a, rel=1e-14)
    assert s["wrand"] == pytest.approx(nr * a2, rel=1e-14)
    assert s["alpha"] == pytest.approx(alpha, rel=1e-14)
    shot = nd * (a * b) ** 2 + alpha ** 2 * nr * (a2 * b2) ** 2
    assert s["shot"] == pytest.approx(shot, rel=1e-13)
    assert not math.isclose(s["shot"], nd * (a * b) ** 2 + alpha * nr * (a2 * b2) ** 2,
                            rel_tol=1e-6)                       # alpha^2, not alpha
    # normalisation from the randoms: alpha sum_r w_comp w_fkp^2 n
    assert s["norm"] == pytest.approx(alpha * nr * a2 * b2 * b2 * c2, rel=1e-13)
    # randoms without n(z): from the data
    _, _, s = oracle.survey_sums(one(nd, a), one(nd, b), one(nd, c),
                                 one(nr, a2), one(nr, b2), one(nr, 0.0))
    assert s["norm"] == pytest.approx(nd * a * b * b * c, rel=1e-13)
    # data without n(z): from the randoms again
    _, _, s = oracle.survey_sums(one(nd, a), one(nd, b), one(nd, 0.0),
                                 one(nr, a2), one(nr, b2), one(nr, c2))
    assert s["norm"] == pytest.approx(alpha * nr * a2 * b2 * b2 * c2, rel=1e-13)
    # zero completeness or FKP weights are refused (src/read_cata.c:165-172)
    for bad in ((0.0, b), (a, 0.0)):
        with pytest.raises(ValueError):
            oracle.survey_sums(one(nd, bad[0]), one(nd, bad[1]), one(nd, c),
                               one(nr, a2), one(nr, b2), one(nr, c2))


# ---- the survey estimator end to end --------------------------------------------

@pytest.mark.parametrize("scheme,intlace", [(1, False), (2, True), (3, True)])
def test_run_survey_matches_brute_force(oracle, scheme, intlace):
    Ng, dk, poles = 8, 0.004, [0, 2, 4]
    d, r, w = _survey_catalogue(150, 1200, 40 + scheme)
    d[:, 3], r[:, 3], sums = oracle.survey_sums(*w)
    conf = oracle.make_conf(gsize=Ng, assign=scheme, intlace=intlace, issim=False, kbin=dk,
                            poles=poles)
    ref = oracle.run(conf, [d], [r], wsums=[sums])
    assert ref["err"] == 0
    cnt, P = _survey_brute(d, r, sums, Ng, scheme, intlace, poles, dk, ref)
    assert np.array_equal(ref["cnt"].astype(np.int64), cnt.astype(np.int64))
    scale = np.abs(P[0]) + sums["shot"] / sums["norm"]
    has = cnt > 0
    for li in range(len(poles)):
        err = np.abs(ref["pl"][0][li] - P[li])[has] / scale[has]
        assert err.max() < 1e-10, (poles[li], err.max())
    # the l > 0 terms are real signal here, not rounding: a 2 pi / norm
    # normalisation (the cross-spectrum factor) would be off by half
    assert np.abs(P[1][has]).max() > 1e-3 * scale[has].max()


def test_run_survey_first_pole_not_zero(oracle):
    """poles = [2, 4]: the reference skips the first pole of a survey when it
    is not 0 (src/multipole.c:1229-1245), so P_2 stays 0 and P_4 is the
    ordinary l = 4 estimate."""
    Ng, dk = 8, 0.004
    d, r, w = _survey_catalogue(120, 900, 9)
    d[:, 3], r[:, 3], sums = oracle.survey_sums(*w)
    conf = oracle.make_conf(gsize=Ng, assign=1, intlace=False, issim=False, kbin=dk,
                            poles=[2, 4])
    ref = oracle.run(conf, [d], [r], wsums=[sums])
    assert ref["err"] == 0
    assert np.all(ref["pl"][0][0] == 0)
    _, P = _survey_brute(d, r, sums, Ng, 1, False, [4], dk, ref)
    scale = np.abs(P[0]).max()
    assert np.allclose(ref["pl"][0][1], P[0], rtol=0, atol=1e-10 * scale)


def test_run_survey_randoms_equal_data(oracle):
    """Randoms identical to the data with alpha = 1: the field vanishes (up to
    the order of the oracle's threaded deposit sums, ~1e-16 of a cell), so
    P_0 = -shot / norm in every populated bin and P_2 = P_4 = 0."""
    Ng, dk = 16, 0.005
    d, _, _ = _survey_catalogue(500, 1, 3)
    sums = {"wdata": 500.0, "wrand": 500.0, "alpha": 1.0, "shot": 812.5, "norm": 0.37}
    conf = oracle.make_conf(gsize=Ng, assign=2, intlace=True, issim=False, kbin=dk)
    ref = oracle.run(conf, [d], [d.copy()], wsums=[sums])
    has = ref["cnt"] > 0
    assert np.all(ref["pl"][0][0][has] == -sums["shot"] / sums["norm"])
    assert np.all(np.abs(ref["pl"][0][1:][:, has]) < 1e-20 * sums["shot"] / sums["norm"])


def test_cross_of_a_catalogue_with_itself(oracle):
    """Two identical catalogues: the cross spectrum carries no shot noise, so
    boxes give cross = auto + (2l+1) shot lcnt / cnt and surveys cross(l=0) =
    auto + shot / norm; the survey l > 0 cross sums both orderings times
    2 pi, which equals the auto term's 4 pi."""
    L, dk, poles = 500.0, 0.02, [0, 2, 4]
    d = _particles(3000, [L] * 3, 8, w=True)
    conf = oracle.make_conf(gsize=16, assign=3, intlace=True, bsize=L, kbin=dk, ndata=2,
                            isauto=[True, True], iscross=True, poles=poles)
    ref = oracle.run(conf, [d, d.copy()])
    has = ref["cnt"] > 0
    shot = ref["shot"][0]
    for li, ell in enumerate(poles):
        want = ref["pl"][0][li] + (2 * ell + 1) * shot * ref["lcnt"][li] / np.maximum(ref["cnt"], 1)
        scale = np.abs(ref["pl"][0][0]) + shot
        assert np.all(np.abs(ref["xpl"][li] - want)[has] <= 1e-12 * scale[has]), ell
    dd, rr, w = _survey_catalogue(200, 1500, 11)
    dd[:, 3], rr[:, 3], sums = oracle.survey_sums(*w)
    conf = oracle.make_conf(gsize=8, assign=1, intlace=False, issim=False, kbin=0.004, ndata=2,
                            isauto=[True, True], iscross=True, poles=poles)
    ref = oracle.run(conf, [dd, dd.copy()], [rr, rr.copy()], wsums=[sums, sums])
    has = ref["cnt"] > 0
    scale = np.abs(ref["pl"][0][0]) + sums["shot"] / sums["norm"]
    d0 = ref["xpl"][0] - (ref["pl"][0][0] + sums["shot"] / sums["norm"])
    assert np.all(np.abs(d0[has]) <= 1e-12 * scale[has])
    for li in (1, 2):
        assert np.all(np.abs(ref["xpl"][li] - ref["pl"][0][li])[has] <= 1e-12 * scale[has])


# ---- interlaced boxes and the periodic shift ------------------------------------

def test_shift_cat_boundary(oracle):
    """x >= smin + L moves down by L (src/genr_mesh.c:590-602): exactly at the
    edge it moves, one ulp below it stays."""
    L = np.array([100.0, 120.0, 80.0])
    smin = np.array([-0.5, -0.25, -1.0])
    edge = smin + L
    d = np.array([[edge[0], np.nextafter(edge[1], -np.inf), edge[2], 1.0],
                  [np.nextafter(edge[0], -np.inf), edge[1], 3.0, 1.0]])
    got = d.copy()
    oracle.shift_cat(got, smin, L)
    assert got[0, 0] == edge[0] - L[0] and got[0, 1] == d[0, 1] and got[0, 2] == edge[2] - L[2]
    assert got[1, 0] == d[1, 0] and got[1, 1] == edge[1] - L[1] and got[1, 2] == 3.0


@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_shifted_grid_deposit_is_periodic(oracle, scheme):
    """shift_cat + assignment on the shifted grid = the periodic window on
    that grid, also for particles exactly at smin + L, whose grid coordinate
    would be N_g without the shift."""
    Ng, L = 8, 80.0
    H = L / Ng
    smin = np.array([-0.5 * H] * 3)
    d = _particles(200, [L] * 3, 60 + scheme, w=True)
    d[:5, 0] = L - 0.5 * H                 # exactly smin + L
    d[5:9, 1] = np.nextafter(L - 0.5 * H, -np.inf)
    s = d.copy()
    oracle.shift_cat(s, smin, [L] * 3)
    got = oracle.assign(scheme, s, Ng, smin, [L] * 3)
    want = brute_mesh(scheme, d, Ng, smin, [L] * 3)
    assert np.allclose(got, want, atol=1e-12)


def _numpy_sim_interlaced(data, Ng, L, dk, poles, scheme):
    """Interlaced box, any scheme: periodic windows on both grids, numpy FFTs,
    the Fourier-space combination, the alias-corrected Legendre sums and the
    box normalisation (shot = V / W, norm = W^2 / V) from the definitions."""
    H = L / Ng
    FA = np.fft.rfftn(brute_mesh(scheme, data, Ng, [0, 0, 0], [L] * 3))
    FB = np.fft.rfftn(brute_mesh(scheme, data, Ng, [-0.5 * H] * 3, [L] * 3))
    F = _interlace_half(FA, FB)
    k2, mult, (kx, ky, kz) = _half_grid(Ng, [L] * 3)
    kmod = np.sqrt(k2)
    nb = int(round(math.pi * Ng / L / dk))
    if dk * nb > math.pi * Ng / L:
        nb -= 1
    kedge = dk * np.arange(nb + 1)
    b, ok = _lin_bins(kmod, kedge, dk)
    cnt = np.bincount(b[ok], weights=mult[ok], minlength=nb)
    ok &= k2 > 0
    mu = np.divide(kz, kmod, out=np.zeros_like(kmod), where=k2 > 0)
    alias = _alias_table(Ng, [L] * 3, True, scheme)
    W = data[:, 3].sum()
    V = L ** 3
    shot, norm = V / W, W * W / V
    out = []
    for ell in poles:
        leg = sps.eval_legendre(ell, mu[ok]) * mult[ok]
        s = np.bincount(b[ok], weights=(np.abs(F) ** 2 * alias)[ok] * leg, minlength=nb)
        lc = np.bincount(b[ok], weights=leg, minlength=nb)
        out.append((s / (norm * cnt) - shot * lc / cnt) * (2 * ell + 1))
    return cnt, np.array(out)


@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
@pytest.mark.parametrize("Ng", [16, 12, 20])
def test_run_sim_interlaced_matches_numpy(oracle, scheme, Ng):
    """Also on grids of 3 * 2^2 and 5 * 2^2 cells, whose transforms the
    oracle takes as direct DFTs."""
    L, dk, poles = 400.0, 0.02, [0, 2, 4]
    H = L / Ng
    d = _particles(800, [L] * 3, 70 + scheme, w=True)
    d[:20, 2] = L - 0.5 * H               # on the shifted grid's edge
    d[20:40, 0] = L - 0.5 * H
    conf = oracle.make_conf(gsize=Ng, assign=scheme, intlace=True, bsize=L, kbin=dk,
                            poles=poles)
    ref = oracle.run(conf, [d])
    assert ref["err"] == 0
    cnt, P = _numpy_sim_interlaced(d, Ng, L, dk, poles, scheme)
    assert np.array_equal(ref["cnt"].astype(np.int64), cnt.astype(np.int64))
    scale = np.abs(P[0]) + L ** 3 / d[:, 3].sum()
    has = cnt > 0
    for li in range(len(poles)):
        err = np.abs(ref["pl"][0][li] - P[li])[has] / scale[has]
        assert err.max() < 1e-10, (poles[li], err.max())


# ---- log bins (src/multipole.c:673-754, 911-1031, rescale :1267-1273) -----------

def test_log_bins_full_grid(oracle):
    Ng, L = 16, 600.0
    kmin, dk = math.log10(0.02), 0.05
    poles = [0, 2, 4]
    d = _particles(600, [L] * 3, 91, w=True)
    conf = oracle.make_conf(gsize=Ng, assign=0, intlace=False, bsize=L, kbin=dk, kmin=kmin,
                            logscale=True, poles=poles)
    ref = oracle.run(conf, [d])
    assert ref["err"] == 0
    kny = math.log10(math.pi * Ng / L)
    nb = int(round((kny - kmin) / dk))
    if kmin + dk * nb > kny:
        nb -= 1
    edges = kmin + dk * np.arange(nb + 1)
    # numpy log binning of every mode of the full grid (NGP: alias 1)
    F = np.fft.fftn(brute_mesh(0, d, Ng, [0, 0, 0], [L] * 3))
    s = _sidx(Ng) * 2 * np.pi / L
    kx, ky, kz = s[:, None, None], s[None, :, None], s[None, None, :]
    k2 = (kx * kx + ky * ky) + kz * kz
    with np.errstate(divide="ignore"):
        lk = 0.5 * np.log10(k2)
    b = np.floor((lk - edges[0]) / dk).astype(np.int64)
    ok = (lk >= edges[0]) & (lk < edges[-1]) & (b >= 0) & (b < nb)
    cnt = np.bincount(b[ok], minlength=nb)
    assert np.array_equal(ref["cnt"].astype(np.int64), cnt)
    # mu = k_z / |k| with |k| the length, not the log (src/multipole.c:940,966)
    mu = np.divide(np.broadcast_to(kz, k2.shape), np.sqrt(k2), out=np.zeros(k2.shape),
                   where=k2 > 0)
    W, V = d[:, 3].sum(), L ** 3
    shot, norm = V / W, W * W / V
    for li, ell in enumerate(poles):
        leg = sps.eval_legendre(ell, mu[ok])
        sm = np.bincount(b[ok], weights=np.abs(F[ok]) ** 2 * leg, minlength=nb)
        lc = np.bincount(b[ok], weights=leg, minlength=nb)
        with np.errstate(invalid="ignore", divide="ignore"):
            want = (sm / (norm * cnt) - shot * lc / cnt) * (2 * ell + 1)
        has = cnt > 0
        assert np.allclose(ref["pl"][0][li][has], want[has], rtol=1e-10,
                           atol=1e-10 * np.abs(want[has]).max()), ell
    # the rescale as the reference writes it: k = 10^centre, then km and the
    # lower edges are 10^k (exponentiated twice), the top edge 10^edge
    cen = (edges[:-1] + edges[1:]) / 2
    assert np.allclose(ref["k"], 10 ** cen, rtol=1e-14)
    assert np.allclose(ref["km"], 10 ** (10 ** cen), rtol=1e-14)
    assert np.allclose(ref["kedge"][:-1], 10 ** (10 ** cen), rtol=1e-14)
    assert ref["kedge"][-1] == pytest.approx(10 ** edges[-1], rel=1e-14)
