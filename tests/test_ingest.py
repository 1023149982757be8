"""Binary catalogue format written by the library (host only, no GPU)."""
import numpy as np
import pytest

import powspec_amd as P


def test_cat_write_layout(tmp_path):
    rng = np.random.default_rng(3)
    n = 1001
    cols = {"x": rng.uniform(0, 1, n), "y": rng.uniform(0, 1, n), "z": rng.uniform(0, 1, n),
            "w_fkp": rng.uniform(0.5, 1, n)}
    path = tmp_path / "c.pscat"
    P.cat_write(path, cols, np.float32)
    raw = path.read_bytes()
    h = P.CatHeader.from_buffer_copy(raw[:64])
    assert h.magic == P.CAT_MAGIC and h.n == n and h.dtype == 4
    assert h.cols == P.COL_X | P.COL_Y | P.COL_Z | P.COL_WFKP
    body = np.frombuffer(raw[64:], np.float32)
    assert len(body) == 4 * n
    for k, name in enumerate(("x", "y", "z", "w_fkp")):
        assert np.array_equal(body[k * n:(k + 1) * n], cols[name].astype(np.float32))


def test_cat_write_rejects_missing_xyz(tmp_path):
    with pytest.raises(P.PshError):
        P.cat_write(tmp_path / "bad.pscat", {"x": np.zeros(3), "y": np.zeros(3)})


def test_ingest_needs_device():
    if P.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(P.PshError) as e:
        P.Ingest()
    assert e.value.code == P.ERR_DEVICE


def test_parse_double_is_correctly_rounded():
    """The decimal conversion of the ASCII ingest against CPython's float()
    (correctly rounded, like glibc strtod used by the reference's sscanf)."""
    rng = np.random.default_rng(11)
    toks = ["0", "-0", "0.0", ".5", "5.", "1e22", "1e23", "9007199254740993",
            "2.2250738585072014e-200", "123456789012345678901234567890", "0.1",
            "0.30000000000000004", "1E+05", "+3.5e-007", "4.9406564584124654e-280",
            "1018344890163455655936", "16020078902493.7041015625"]
    for _ in range(20000):
        k = rng.integers(0, 5)
        m = rng.random() * 10.0 ** int(rng.integers(-40, 40))
        if k == 0:
            toks.append(repr(m))
        elif k == 1:
            toks.append(f"{m:.{int(rng.integers(0, 25))}e}")
        elif k == 2:
            toks.append(f"{rng.random() * 2000:.{int(rng.integers(0, 18))}f}")
        elif k == 3:
            toks.append("".join(rng.choice(list("0123456789"), int(rng.integers(15, 45)))) )
        else:   # exact binary midpoints written out in full
            x = float(rng.integers(2**52, 2**53))
            toks.append(f"{(x + 0.5) * 2.0 ** int(rng.integers(-20, 20)):.40g}")
    for t in toks:
        got, ref = P.parse_double(t), float(t)
        assert got == ref and np.signbit(got) == np.signbit(ref), t


def test_parse_double_rejects():
    for t in ["", "+", ".", "e5", "1e", "1e+", "1.2.3", "nan", "inf", "0x1p3", "1,5", "--1",
              "1e400"]:
        with pytest.raises(P.PshError):
            P.parse_double(t)
