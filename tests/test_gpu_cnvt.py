"""Survey coordinate conversion on the GPU (psh_cnvt_device / psh_cnvt_coord)
against the FP64 oracle (src/cnvt_coord.c): the same Gauss-Legendre order,
positions within 1e-12 relative; spline interpolation; errors."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _sky(n, seed, zlo=0.05, zhi=1.6):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(0, 360, n), np.degrees(np.arcsin(rng.uniform(-1, 1, n))),
                            rng.uniform(zlo, zhi, n), rng.uniform(0.5, 1.5, n)])


@pytest.mark.parametrize("om,ol,w,err", [(0.31, None, -1.0, 1e-8), (0.3, 0.65, -0.9, 1e-11),
                                         (1.0, 0.0, -1.0, 1e-6)])
def test_integration(psh, om, ol, w, err):
    d = _sky(200000, 5)
    cv = psh.make_cnvt(omega_m=om, omega_l=ol, eos_w=w, ecdst=err)
    dev = psh.DeviceArray.from_numpy(d)
    order = psh.cnvt_device(cv, [(dev, len(d))])
    ref_order = O.cmvdist_order(cv.omega_m, cv.omega_l, cv.omega_k, O.eos_widx(w), err,
                                d[:, 2].min(), d[:, 2].max())
    assert order == ref_order
    got = dev.to_numpy((len(d), 4))
    ref = O.cnvt(d, om=cv.omega_m, ol=cv.omega_l, ok=cv.omega_k, eos_w=w, order=order)
    r = np.linalg.norm(ref[:, :3], axis=1)
    assert np.all(np.abs(got[:, :3] - ref[:, :3]) <= 1e-12 * r[:, None])
    assert np.array_equal(got[:, 3], d[:, 3])


def test_interpolation(psh):
    zs = np.linspace(0.0, 2.0, 301)
    ds = np.array([O.cmvdist(0.31, 0.69, 0.0, 0.0, 32, z) for z in zs])
    d = _sky(100000, 6, 0.0, 2.2)           # some redshifts beyond the samples
    d[0, 2] = 0.0
    d[1, 2] = 2.0                            # the upper end is excluded
    cv = psh.make_cnvt(samples=(zs, ds))
    dev = psh.DeviceArray.from_numpy(d)
    assert psh.cnvt_device(cv, [(dev, len(d))]) == 0
    got = dev.to_numpy((len(d), 4))
    ref = O.cnvt(d, zs=zs, ds=ds)
    fin = np.isfinite(ref[:, 2]) & np.isfinite(ref[:, 0])
    assert np.all(~np.isfinite(got[~fin, 0]) | ~np.isfinite(got[~fin, 2]))
    r = np.linalg.norm(ref[fin, :3], axis=1)
    assert np.all(np.abs(got[fin, :3] - ref[fin, :3]) <= 1e-12 * (r[:, None] + 1))


def test_dropin_and_errors(psh):
    d = _sky(30000, 7)
    r = _sky(90000, 8)
    cat = psh.Cata()
    keep = [d.copy(), r.copy()]   # converted in place by the library
    dp = (C.c_void_p * 2)(keep[0].ctypes.data, None)
    rp = (C.c_void_p * 2)(keep[1].ctypes.data, None)
    nd = (C.c_size_t * 2)(len(d), 0)
    nr = (C.c_size_t * 2)(len(r), 0)
    cat.num = 1
    cat.data, cat.rand = C.cast(dp, C.c_void_p), C.cast(rp, C.c_void_p)
    cat.ndata, cat.nrand = C.cast(nd, C.c_void_p), C.cast(nr, C.c_void_p)
    cv = psh.make_cnvt(omega_m=0.31, dcnvt=(True, False), rcnvt=(True, False))
    assert psh.load().psh_cnvt_coord(C.byref(cv), C.byref(cat)) == 0
    both = np.concatenate([d, r])
    order = O.cmvdist_order(0.31, 0.69, 0.0, 0.0, 1e-8, both[:, 2].min(), both[:, 2].max())
    for got, src in zip(keep, (d, r)):
        ref = O.cnvt(src, order=order)
        rr = np.linalg.norm(ref[:, :3], axis=1)
        assert np.all(np.abs(got[:, :3] - ref[:, :3]) <= 1e-12 * rr[:, None])
    bad = _sky(1000, 9)
    bad[10, 2] = -0.01
    dev = psh.DeviceArray.from_numpy(bad)
    with pytest.raises(psh.PshError) as e:
        psh.cnvt_device(psh.make_cnvt(), [(dev, len(bad))])
    assert e.value.code == psh.ERR_CNVT
