"""Pins of the oracle's math building blocks (CPU only).

Each function of oracle/psoracle.c is checked against something other than a
retyping of its own formula: the reference's own header-only math compiled
as-is (oracle/_ref, when /root/reference is present), scipy, numpy's FFT,
direct DFTs, closed forms and invariants.
"""
import math

import numpy as np
import pytest
import scipy.special as sps


# ---- FFT: against numpy (pocketfft) and a direct DFT -------------------------

# 6, 12, 20, 24, 40, 48, 50, 18: the odd-factor step (odd parts 3, 5, 9, 25);
# 9, 14: the direct DFT
@pytest.mark.parametrize("Ng", [4, 6, 8, 9, 12, 14, 16, 18, 20, 24, 32, 40, 48, 50])
def test_fft_r2c_matches_numpy(oracle, Ng):
    rng = np.random.default_rng(Ng)
    x = rng.standard_normal((Ng, Ng, Ng))
    F = oracle.fft_r2c(x)
    R = np.fft.rfftn(x)
    assert F.shape == (Ng, Ng, Ng // 2 + 1)
    assert np.max(np.abs(F - R)) <= 1e-13 * np.max(np.abs(R))


def test_fft_r2c_direct_dft(oracle):
    Ng = 4
    rng = np.random.default_rng(3)
    x = rng.standard_normal((Ng, Ng, Ng))
    n = np.arange(Ng)
    W = np.exp(-2j * np.pi * np.outer(n, n) / Ng)
    D = np.einsum("ai,bj,ck,ijk->abc", W, W, W, x)[:, :, : Ng // 2 + 1]
    assert np.allclose(oracle.fft_r2c(x), D, atol=1e-12)


def test_fft_delta_and_plane_wave(oracle):
    Ng = 16
    x = np.zeros((Ng, Ng, Ng))
    x[0, 0, 0] = 1.0
    assert np.allclose(oracle.fft_r2c(x), 1.0)          # flat spectrum
    i, j, k = np.meshgrid(*(np.arange(Ng),) * 3, indexing="ij")
    pw = np.cos(2 * np.pi * (3 * i + 2 * j + 5 * k) / Ng)
    F = oracle.fft_r2c(pw)
    nz = np.argwhere(np.abs(F) > 1e-8)
    # cos splits into the +/- mode; only (3,2,5) lies in the stored half
    assert nz.tolist() == [[3, 2, 5]]
    assert abs(F[3, 2, 5] - Ng ** 3 / 2) < 1e-8


def test_fft_parseval_and_c2r(oracle):
    Ng = 16
    rng = np.random.default_rng(5)
    x = rng.standard_normal((Ng, Ng, Ng))
    F = oracle.fft_r2c(x)
    w = np.full(Ng // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0
    assert np.isclose((np.abs(F) ** 2 * w).sum() / Ng ** 3, (x * x).sum(), rtol=1e-12)
    assert np.allclose(oracle.fft_c2r(F, Ng) / Ng ** 3, x, atol=1e-13)


@pytest.mark.parametrize("Ng", [8, 12, 20])
def test_fft_c2r_matches_numpy_on_hermitian_input(oracle, Ng):
    rng = np.random.default_rng(9)
    F = np.fft.rfftn(rng.standard_normal((Ng, Ng, Ng)))
    assert np.allclose(oracle.fft_c2r(F, Ng), np.fft.irfftn(F, s=(Ng,) * 3, axes=(0, 1, 2)) * Ng ** 3,
                       atol=1e-11)


def test_fft_c2r_ignores_imag_of_dc_and_nyquist_rows(oracle):
    # c2r along z ignores Im of the kz=0 and kz=Ng/2 entries after the x,y
    # inverse transforms; a purely imaginary z-DC/Nyquist input gives 0 output
    # when it is Hermitian in (kx, ky).
    Ng = 8
    F = np.zeros((Ng, Ng, Ng // 2 + 1), complex)
    F[0, 0, 0] = 1j
    F[0, 0, Ng // 2] = 2j
    assert np.allclose(oracle.fft_c2r(F, Ng), 0.0, atol=1e-14)


# ---- Legendre and spherical harmonics ---------------------------------------

def test_legpoly_vs_scipy(oracle):
    for ell in range(7):
        for x in np.linspace(-1, 1, 41):
            assert abs(oracle.legpoly(ell, x) - sps.eval_legendre(ell, x)) < 1e-14
    assert oracle.legpoly(7, 0.3) == 0.0


def test_legpoly_vs_reference_header(oracle):
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("reference math headers not built (oracle/_ref)")
    for ell in range(7):
        for x in np.linspace(-1, 1, 57):
            assert abs(oracle.legpoly(ell, x) - R.ref_legpoly(ell, x)) <= 4e-16 * 20


def _angles(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    rxy = np.hypot(v[:, 0], v[:, 1])
    return v, v[:, 2], rxy, v[:, 0] / rxy, v[:, 1] / rxy


def test_ylm_vs_reference_header(oracle):
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("reference math headers not built (oracle/_ref)")
    _, ct, st, cp, sp = _angles(200, 1)
    for ell in range(1, 7):
        for m in range(-ell, ell + 1):
            for q in range(len(ct)):
                a = oracle.ylm(ell, m, ct[q], st[q], cp[q], sp[q])
                b = R.ref_ylm(ell, m, ct[q], st[q], cp[q], sp[q])
                assert abs(a - b) <= 1e-13 * max(1.0, abs(b)), (ell, m, a, b)


def test_ylm_addition_theorem(oracle):
    # sum_m Y_lm(a) Y_lm(b) = (2l+1)/(4 pi) P_l(a.b): pins every normalisation
    # constant and the cos/sin(m phi) assignment.
    va, cta, sta, cpa, spa = _angles(20, 2)
    vb, ctb, stb, cpb, spb = _angles(20, 3)
    for ell in range(1, 7):
        for q in range(20):
            s = sum(oracle.ylm(ell, m, cta[q], sta[q], cpa[q], spa[q]) *
                    oracle.ylm(ell, m, ctb[q], stb[q], cpb[q], spb[q])
                    for m in range(-ell, ell + 1))
            rhs = (2 * ell + 1) / (4 * math.pi) * sps.eval_legendre(ell, va[q] @ vb[q])
            assert abs(s - rhs) < 1e-13


def test_ylm_vs_scipy_real_form(oracle):
    # real form without Condon-Shortley phase:
    #   m>0: sqrt2 (-1)^m Re Y_l^m ; m<0: sqrt2 (-1)^m Im Y_l^|m| ; m=0: Y_l^0
    _, ct, st, cp, sp = _angles(30, 4)
    theta = np.arccos(ct)
    phi = np.arctan2(sp, cp)
    for ell in range(1, 7):
        for m in range(-ell, ell + 1):
            Y = sps.sph_harm_y(ell, abs(m), theta, phi)
            if m > 0:
                ref = math.sqrt(2) * (-1) ** m * Y.real
            elif m < 0:
                ref = math.sqrt(2) * (-1) ** m * Y.imag
            else:
                ref = Y.real
            got = np.array([oracle.ylm(ell, m, ct[q], st[q], cp[q], sp[q])
                            for q in range(len(ct))])
            assert np.allclose(got, ref, atol=1e-13), (ell, m)


# ---- alias correction ---------------------------------------------------------

def test_alias_corr_closed_forms(oracle):
    x = 0.7
    s2 = math.sin(x) ** 2
    for p, scheme in enumerate([0, 1, 2, 3], start=1):
        assert math.isclose(oracle.alias_corr(True, scheme, x), (x / math.sin(x)) ** (2 * p),
                            rel_tol=1e-15)
    assert oracle.alias_corr(False, 0, x) == 1.0
    assert math.isclose(oracle.alias_corr(False, 1, x), 1 / (1 - 2 / 3 * s2), rel_tol=1e-15)
    assert math.isclose(oracle.alias_corr(False, 2, x),
                        1 / (1 - s2 + 2 / 15 * s2 * s2), rel_tol=1e-15)
    assert math.isclose(oracle.alias_corr(False, 3, x),
                        1 / (1 - 4 / 3 * s2 + 2 / 5 * s2 ** 2 - 4 / 315 * s2 ** 3),
                        rel_tol=1e-15)
    # the corrections are 1 at k=0
    for scheme in range(4):
        for il in (True, False):
            assert oracle.alias_corr(il, scheme, 0.0) == 1.0


def test_alias_corr_negative_frequencies_use_series(oracle):
    # The reference takes the Taylor branch for every pik <= 0.01, so negative
    # frequencies get the truncated series (src/multipole.c:50-58).
    x = -1.4
    t = x * x
    series = 1 + t / 6 + 7 * t * t / 360 + 31 * t ** 3 / 15120
    assert math.isclose(oracle.alias_corr(True, 1, x), series ** 4, rel_tol=1e-15)
    assert abs(oracle.alias_corr(True, 1, x) - (x / math.sin(x)) ** 4) > 1e-3
    sin_series = (1 - t / 6 + t * t / 120 - t ** 3 / 5040) * x
    assert math.isclose(oracle.alias_corr(False, 1, x),
                        1 / (1 - 2 / 3 * sin_series ** 2), rel_tol=1e-15)


# ---- k bins ---------------------------------------------------------------------

def test_kbins_closed_form(oracle):
    # N_g = 256, L = 1000 geometry: kNy = pi*256/1000 = 0.80425, dk=0.01 -> round(80.4)=80 bins
    nb, kedge, kc = oracle.kbins(256, [1000.0] * 3, False, 0.0, -1.0, 0.01)
    assert nb == 80
    assert np.allclose(kedge, 0.01 * np.arange(81))
    assert np.allclose(kc, 0.01 * np.arange(80) + 0.005)
    # N_g = 1024, L = 1000 geometry: kNy = 3.21699, dk = 0.005 -> round(643.4) = 643
    nb, _, _ = oracle.kbins(1024, [1000.0] * 3, False, 0.0, -1.0, 0.005)
    assert nb == 643
    # trimming: kNy = pi*16/100 = 0.50265; dk 0.2 -> round(2.51)=3 but 0.6>kNy -> 2
    nb, _, _ = oracle.kbins(16, [100.0] * 3, False, 0.0, -1.0, 0.2)
    assert nb == 2
    # kmax below kNy wins
    nb, _, _ = oracle.kbins(256, [1000.0] * 3, False, 0.0, 0.3, 0.01)
    assert nb == 30
    # non-cubic: Nyquist of the largest side
    nb, _, _ = oracle.kbins(64, [500.0, 800.0, 600.0], False, 0.0, -1.0, 0.01)
    assert nb == round(math.pi * 64 / 800 / 0.01) - (0.01 * round(math.pi * 64 / 800 / 0.01) > math.pi * 64 / 800)
    # not enough bins
    nb, _, _ = oracle.kbins(4, [1000.0] * 3, False, 0.0, -1.0, 0.5)
    assert nb < 0
