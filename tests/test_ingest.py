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
