"""Pins of the oracle's survey coordinate conversion (src/cnvt_coord.c):
Gauss-Legendre rules vs numpy, distances vs closed forms, the order search,
the natural spline vs scipy, and the (RA, Dec) -> (x, y, z) mapping."""
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import oracle as O

C_H0 = 299792.458 * 0.01   # c / (100 km/s/Mpc) in Mpc/h


@pytest.mark.parametrize("n", [4, 5, 8, 13, 20, 31, 32])
def test_legauss_matches_numpy(n):
    x, w, w0 = O.legauss(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    pos = xr > 1e-14
    order = np.argsort(-xr[pos])
    assert np.allclose(x, xr[pos][order], rtol=0, atol=2e-15)
    assert np.allclose(w, wr[pos][order], rtol=1e-13)
    if n & 1:
        assert math.isclose(w0, wr[np.abs(xr) < 1e-14][0], rel_tol=1e-13)
    assert math.isclose(2 * w.sum() + w0, 2.0, rel_tol=1e-14)


@pytest.mark.parametrize("z", [0.1, 0.5, 1.0, 2.5])
def test_einstein_de_sitter_distance(z):
    """Om = 1: D = 2 c/H0 (1 - 1/sqrt(1+z))."""
    got = O.cmvdist(1.0, 0.0, 0.0, 0.0, 32, z)
    assert math.isclose(got, 2 * C_H0 * (1 - 1 / math.sqrt(1 + z)), rel_tol=1e-13)


@pytest.mark.parametrize("z", [0.2, 1.3])
def test_de_sitter_distance_and_w_index(z):
    """OL = 1, w = -1: D = c/H0 z (constant integrand: any order is exact);
    OL = 1 with w = 0 behaves as matter: the widx = 3(1+w) power."""
    assert math.isclose(O.cmvdist(0.0, 1.0, 0.0, 0.0, 4, z), C_H0 * z, rel_tol=1e-14)
    got = O.cmvdist(0.0, 1.0, 0.0, O.eos_widx(0.0), 32, z)
    assert math.isclose(got, 2 * C_H0 * (1 - 1 / math.sqrt(1 + z)), rel_tol=1e-13)


def test_order_search():
    """The chosen order meets the requested precision at the worst redshift,
    grows with the precision, and a constant integrand needs the minimum."""
    om, ol, err = 0.31, 0.69, 1e-10
    n = O.cmvdist_order(om, ol, 0.0, 0.0, err, 0.01, 3.0)
    assert 4 <= n <= 32
    exact = O.cmvdist(om, ol, 0.0, 0.0, 32, 3.0)
    assert abs(O.cmvdist(om, ol, 0.0, 0.0, n, 3.0) - exact) <= 10 * err * exact
    assert O.cmvdist_order(om, ol, 0.0, 0.0, 1e-4, 0.01, 3.0) <= n
    assert O.cmvdist_order(0.0, 1.0, 0.0, 0.0, 1e-12, 0.0, 2.0) == 5
    assert O.cmvdist_order(om, ol, 0.0, 0.0, 1e-30, 0.0, 50.0) == -1   # never converges


def test_spline_matches_scipy_natural():
    rng = np.random.default_rng(3)
    zs = np.sort(np.concatenate([[0.0], rng.uniform(0, 2, 40), [2.0]]))
    ds = np.sin(zs) * 3000 + zs ** 3
    ypp = O.spline_ypp(zs, ds)
    cs = CubicSpline(zs, ds, bc_type="natural")
    assert np.allclose(ypp, cs(zs, 2), rtol=1e-9, atol=1e-9 * np.abs(ypp).max())
    for z in np.concatenate([rng.uniform(0, 2, 200), zs[:-1]]):
        assert math.isclose(O.spline_dist(zs, ds, ypp, z), float(cs(z)), rel_tol=1e-12,
                            abs_tol=1e-9)


def test_spline_linear_and_range():
    zs = np.linspace(0.0, 1.0, 11)
    ds = 1000.0 * zs + 5.0
    ypp = O.spline_ypp(zs, ds)
    assert np.allclose(ypp, 0.0, atol=1e-9)
    assert math.isclose(O.spline_dist(zs, ds, ypp, 0.37), 375.0, rel_tol=1e-14)
    assert O.spline_dist(zs, ds, ypp, -0.1) == math.inf
    assert O.spline_dist(zs, ds, ypp, 1.0) == math.inf       # upper end excluded
    assert O.spline_dist(zs, ds, ypp, 0.0) == 5.0


def test_cnvt_geometry():
    """Positions lie at the comoving distance, on the (RA, Dec) direction."""
    rng = np.random.default_rng(8)
    d = np.column_stack([rng.uniform(0, 360, 500), rng.uniform(-90, 90, 500),
                         rng.uniform(0.1, 1.5, 500), np.ones(500)])
    out = O.cnvt(d, order=32)
    r = np.linalg.norm(out[:, :3], axis=1)
    ref = np.array([O.cmvdist(0.31, 0.69, 0.0, 0.0, 32, z) for z in d[:, 2]])
    assert np.allclose(r, ref, rtol=1e-13)
    ra, dec = np.radians(d[:, 0]), np.radians(d[:, 1])
    assert np.allclose(out[:, 2] / r, np.sin(dec), atol=1e-13)
    assert np.allclose(np.arctan2(out[:, 1], out[:, 0]) % (2 * np.pi), ra % (2 * np.pi),
                       atol=1e-9)
    assert np.array_equal(out[:, 3], d[:, 3])
