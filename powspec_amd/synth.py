"""Seeded synthetic catalogues (host side, numpy), drawing the same
Philox4x32-10 streams as the GPU generators (powspec_amd/csrc/k_gen.hip,
psh_gen_*): counter = (index_lo, index_hi, stream, draw), key = (seed_lo,
seed_hi), four 32-bit words per counter (Salmon et al. 2011).

* uniform_box: positions uniform in [0, L)^3, unit weights -- bit-identical
  to psh_gen_uniform_box.
* clustered_box: Gaussian field with P(k) = A (k/0.1)^n exp(-(k/kc)^2) on a
  generator grid (white noise, FFT, amplitude, inverse FFT), lognormal
  transform, Poisson sampling at mean density nbar (positions uniform inside
  each cell), Zel'dovich redshift-space shift along z, s_z = z + f Psi_z
  with Psi(k) = i k / k^2 delta_G(k), periodic wrap.  Same streams as
  psh_gen_clustered_box; the host FFTs are FP64, the GPU's FP32.
* sky: survey (RA, Dec, z, w_fkp) records, as psh_gen_sky.

Records are psh_data (x, y, z, w) float64 rows.  Every coordinate is kept
strictly below L so that (x * Ng / L) stays below Ng for every grid size.
"""
from __future__ import annotations

import numpy as np

try:
    import scipy.fft as _fft

    def _rfftn(a, workers):
        return _fft.rfftn(a, workers=workers)

    def _irfftn(a, s, workers):
        return _fft.irfftn(a, s=s, workers=workers)
except ImportError:  # pragma: no cover
    def _rfftn(a, workers):
        return np.fft.rfftn(a)

    def _irfftn(a, s, workers):
        return np.fft.irfftn(a, s=s)

# streams (powspec_amd/csrc/philox.h)
UNIFORM_POS, NOISE, POISSON, PLACE, SKY = 1, 3, 4, 5, 6
_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)
_CHUNK = 1 << 22


def philox4x32_10(c0, c1, c2, c3, k0: int, k1: int):
    """Philox4x32-10 on arrays of counters (any unsigned ints < 2^32)."""
    c0, c1, c2, c3 = (np.asarray(c, np.uint64) & _MASK for c in (c0, c1, c2, c3))
    c0, c1, c2, c3 = np.broadcast_arrays(c0, c1, c2, c3)
    for _ in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        n0 = (p1 >> np.uint64(32)) ^ c1 ^ np.uint64(k0)
        n2 = (p0 >> np.uint64(32)) ^ c3 ^ np.uint64(k1)
        c1 = p1 & _MASK
        c3 = p0 & _MASK
        c0, c2 = n0, n2
        k0 = (k0 + _W0) & 0xFFFFFFFF
        k1 = (k1 + _W1) & 0xFFFFFFFF
    return c0, c1, c2, c3


def draw(seed: int, index, stream: int, d: int):
    idx = np.asarray(index, np.uint64)
    return philox4x32_10(idx & _MASK, idx >> np.uint64(32), stream, d,
                         seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)


def u53(a, b):
    """53-bit uniform double in [0, 1) from two words."""
    return ((a >> np.uint64(5)).astype(np.float64) * 67108864.0 +
            (b >> np.uint64(6)).astype(np.float64)) * (1.0 / 9007199254740992.0)


def _chunks(n):
    for s in range(0, n, _CHUNK):
        yield s, min(n, s + _CHUNK)


def box_top(L):
    """Largest coordinate kept: x * Ng / L stays below Ng for every Ng."""
    return L * (1.0 - 4e-16) - 1e-9 * L


def _guard(pos, L):
    """Wrap into [0, L) and keep pos * Ng / L < Ng for all Ng."""
    pos = np.mod(pos, L)
    return np.minimum(pos, box_top(L))


def uniform_box(n: int, L: float = 1000.0, seed: int = 1) -> np.ndarray:
    Ls = np.broadcast_to(np.asarray(L, float), (3,))
    d = np.empty((n, 4))
    for s, e in _chunks(n):
        a = draw(seed, np.arange(s, e, dtype=np.uint64), UNIFORM_POS, 0)
        b = draw(seed, np.arange(s, e, dtype=np.uint64), UNIFORM_POS, 1)
        d[s:e, 0] = np.minimum(u53(a[0], a[1]) * Ls[0], box_top(Ls[0]))
        d[s:e, 1] = np.minimum(u53(a[2], a[3]) * Ls[1], box_top(Ls[1]))
        d[s:e, 2] = np.minimum(u53(b[0], b[1]) * Ls[2], box_top(Ls[2]))
    d[:, 3] = 1.0
    return d


def normals(seed: int, index, stream: int = NOISE):
    """One Box-Muller normal per counter (draw 0: u1 from words 0-1, u2 from
    words 2-3)."""
    w = draw(seed, index, stream, 0)
    u1, u2 = u53(w[0], w[1]), u53(w[2], w[3])
    return np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * np.pi * u2)


def poisson_inv(lam, u):
    """Poisson deviates by CDF inversion with one uniform each (normal
    approximation with the same uniform above lam = 600)."""
    lam = np.asarray(lam, np.float64)
    u = np.asarray(u, np.float64)
    k = np.zeros(lam.shape, np.int64)
    big = lam > 600
    small = (lam > 0) & ~big
    if big.any():
        from scipy.special import erfinv
        lb = lam[big]
        k[big] = np.maximum(np.floor(lb + np.sqrt(lb) * np.sqrt(2) * erfinv(2 * u[big] - 1)
                                     + 0.5), 0).astype(np.int64)
    idx = np.nonzero(small)[0]
    if idx.size:
        ls, us = lam[idx], u[idx]
        p = np.exp(-ls)
        cdf = p.copy()
        kk = np.zeros(idx.size, np.int64)
        cap = ls + 40 * np.sqrt(ls) + 100
        act = (us >= cdf) & (kk < cap)
        while act.any():
            kk[act] += 1
            p[act] *= ls[act] / kk[act]
            cdf[act] += p[act]
            act = act & (us >= cdf) & (kk < cap)
        k[idx] = kk
    return k


def power_law_cut(k, A=1e4, n=-1.5, kc=0.5):
    with np.errstate(divide="ignore", invalid="ignore"):
        p = A * (k / 0.1) ** n * np.exp(-(k / kc) ** 2)
    p[k == 0] = 0.0
    return p


def gaussian_field(ng: int, L: float, seed: int, pk=power_law_cut, workers=-1):
    """delta_G and the z displacement Psi_z on an ng^3 grid (float32), from
    the white-noise stream of psh_gen_clustered_box (FP64 FFTs here)."""
    ncell = ng ** 3
    w = np.empty(ncell, np.float32)
    for s, e in _chunks(ncell):
        w[s:e] = normals(seed, np.arange(s, e, dtype=np.uint64))
    wk = _rfftn(w.reshape(ng, ng, ng).astype(np.float64), workers)
    del w
    kf = 2 * np.pi / L
    kx = np.fft.fftfreq(ng, d=1.0 / ng) * kf
    kz = np.arange(ng // 2 + 1) * kf
    k2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    amp = np.sqrt(pk(np.sqrt(k2)) * ng ** 3 / L ** 3)
    wk *= amp
    del amp
    delta = _irfftn(wk, (ng, ng, ng), workers).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        fac = np.where(k2 > 0, kz[None, None, :] / k2, 0.0)
    del k2
    psi_z = _irfftn(wk * (1j * fac), (ng, ng, ng), workers).astype(np.float32)
    return delta, psi_z


def clustered_box(ng_gen: int = 512, L: float = 1000.0, nbar: float = 0.02,
                  f: float = 0.8, seed: int = 2, A: float = 1e4, n: float = -1.5,
                  kc: float = 0.5, workers: int = -1, npart: float | None = None,
                  return_counts: bool = False):
    """Clustered periodic box; npart, if given, sets the mean density so that
    the expected particle count is npart."""
    delta, psi_z = gaussian_field(ng_gen, L, seed,
                                  pk=lambda k: power_law_cut(k, A, n, kc), workers=workers)
    ncell = ng_gen ** 3
    dd = delta.ravel().astype(np.float64)
    var = float(np.mean(dd * dd) - np.mean(dd) ** 2)
    cell = L / ng_gen
    mean = nbar * cell ** 3 if not npart else npart / ncell
    counts = np.empty(ncell, np.int64)
    for s, e in _chunks(ncell):
        w = draw(seed, np.arange(s, e, dtype=np.uint64), POISSON, 0)
        counts[s:e] = poisson_inv(mean * np.exp(dd[s:e] - 0.5 * var), u53(w[0], w[1]))
    del dd
    ntot = int(counts.sum())
    out = np.empty((ntot, 4))
    cells = np.repeat(np.arange(ncell, dtype=np.int64), counts)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    slot = np.arange(ntot, dtype=np.int64) - np.repeat(starts, counts)
    key = (cells.astype(np.uint64) << np.uint64(16)) | (slot.astype(np.uint64) & np.uint64(0xFFFF))
    a = draw(seed, key, PLACE, 0)
    b = draw(seed, key, PLACE, 1)
    i, j, k = cells // (ng_gen * ng_gen), (cells // ng_gen) % ng_gen, cells % ng_gen
    out[:, 0] = _guard((i + u53(a[0], a[1])) * cell, L)
    out[:, 1] = _guard((j + u53(a[2], a[3])) * cell, L)
    out[:, 2] = _guard((k + u53(b[0], b[1])) * cell + f * psi_z.ravel()[cells].astype(np.float64), L)
    out[:, 3] = 1.0
    if return_counts:
        return out, counts
    return out


def sky(n: int, seed: int, ra=(0.0, 60.0), dec=(-30.0, 30.0), zmu=0.55, zsigma=0.08,
        zlo=0.4, zhi=0.7, nbar0=3e-4, p0=1e4):
    """(RA, Dec, z, w_fkp) records and n(z) as psh_gen_sky."""
    from scipy.special import erfc, erfinv
    phi = lambda x: 0.5 * erfc(-x / np.sqrt(2))
    s0, s1 = np.sin(np.radians(dec[0])), np.sin(np.radians(dec[1]))
    pa, pb = phi((zlo - zmu) / zsigma), phi((zhi - zmu) / zsigma)
    out = np.empty((n, 4))
    nz = np.empty(n)
    for s, e in _chunks(n):
        idx = np.arange(s, e, dtype=np.uint64)
        a = draw(seed, idx, SKY, 0)
        b = draw(seed, idx, SKY, 1)
        out[s:e, 0] = ra[0] + (ra[1] - ra[0]) * u53(a[0], a[1])
        out[s:e, 1] = np.degrees(np.arcsin(s0 + (s1 - s0) * u53(a[2], a[3])))
        z = zmu + zsigma * np.sqrt(2) * erfinv(2 * (pa + (pb - pa) * u53(b[0], b[1])) - 1)
        nzv = nbar0 * np.exp(-0.5 * ((z - zmu) / zsigma) ** 2)
        out[s:e, 2] = z
        out[s:e, 3] = 1.0 / (1.0 + nzv * p0)
        nz[s:e] = nzv
    return out, nz
