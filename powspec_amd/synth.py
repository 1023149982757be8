"""Seeded synthetic catalogues for tests and benchmarks (host side, numpy).

* uniform_box: positions uniform in [0, L)^3, unit weights.
* clustered_box: Gaussian field with P(k) = A (k/0.1)^n exp(-(k/kc)^2) on a
  generator grid, lognormal transform, Poisson sampling at mean density nbar
  (uniform positions inside each cell), Zel'dovich redshift-space shift along
  z, s_z = z + f Psi_z with Psi(k) = i k / k^2 delta_G(k), periodic wrap.

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


def _guard(pos, L):
    """Wrap into [0, L) and keep pos * Ng / L < Ng for all Ng."""
    pos = np.mod(pos, L)
    top = L * (1.0 - 4e-16) - 1e-9 * L
    return np.minimum(pos, top)


def uniform_box(n: int, L: float = 1000.0, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    d = np.empty((n, 4))
    d[:, :3] = _guard(rng.random((n, 3)) * L, L)
    d[:, 3] = 1.0
    return d


def power_law_cut(k, A=1e4, n=-1.5, kc=0.5):
    with np.errstate(divide="ignore"):
        p = A * (k / 0.1) ** n * np.exp(-(k / kc) ** 2)
    p[k == 0] = 0.0
    return p


def gaussian_field(ng: int, L: float, seed: int, pk=power_law_cut, workers=-1):
    """delta_G and the z displacement Psi_z on an ng^3 grid (float32)."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((ng, ng, ng), dtype=np.float32)
    wk = _rfftn(w, workers)
    del w
    kf = 2 * np.pi / L
    kx = np.fft.fftfreq(ng, d=1.0 / ng).astype(np.float32) * kf
    kz = np.arange(ng // 2 + 1, dtype=np.float32) * kf
    k2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    k = np.sqrt(k2)
    amp = np.sqrt(pk(k.astype(np.float64)) * ng ** 3 / L ** 3).astype(np.float32)
    del k
    wk *= amp
    del amp
    delta = _irfftn(wk, (ng, ng, ng), workers).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        fac = np.where(k2 > 0, kz[None, None, :] / k2, 0.0).astype(np.float32)
    del k2
    psik = wk * (1j * fac)
    del wk, fac
    psi_z = _irfftn(psik, (ng, ng, ng), workers).astype(np.float32)
    return delta, psi_z


def clustered_box(ng_gen: int = 512, L: float = 1000.0, nbar: float = 0.02,
                  f: float = 0.8, seed: int = 2, A: float = 1e4, n: float = -1.5,
                  kc: float = 0.5, workers: int = -1, npart: int | None = None):
    """Clustered periodic box; npart, if given, rescales the mean density so
    that the expected particle count is npart."""
    delta, psi_z = gaussian_field(ng_gen, L, seed,
                                  pk=lambda k: power_law_cut(k, A, n, kc),
                                  workers=workers)
    var = float(np.var(delta, dtype=np.float64))
    rho = np.exp(delta - 0.5 * var, dtype=np.float32)   # 1 + delta_LN
    del delta
    cell = L / ng_gen
    mean = nbar * cell ** 3 if npart is None else npart / ng_gen ** 3
    rng = np.random.default_rng(seed + 7919)
    counts = rng.poisson(rho.ravel() * np.float32(mean))
    del rho
    ntot = int(counts.sum())
    idx = np.repeat(np.arange(counts.size, dtype=np.int64), counts)
    del counts
    out = np.empty((ntot, 4))
    i = idx // (ng_gen * ng_gen)
    j = (idx // ng_gen) % ng_gen
    k = idx % ng_gen
    u = rng.random((ntot, 3))
    out[:, 0] = (i + u[:, 0]) * cell
    out[:, 1] = (j + u[:, 1]) * cell
    out[:, 2] = (k + u[:, 2]) * cell + f * psi_z.ravel()[idx]
    del u, i, j, k, idx
    out[:, :3] = _guard(out[:, :3], L)
    out[:, 3] = 1.0
    return out
