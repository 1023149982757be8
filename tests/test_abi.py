"""The C-ABI boundary on the CPU: every function include/powspec_hip.h
declares is exported by the built library, the ctypes mirror of the structs
has the C layout (gcc on the header), the status codes match, and without a
GPU the product path reports an error instead of computing anything."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import powspec_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "powspec_hip.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(psh_[a-z0-9_]+)\s*\(", src)))


def test_header_functions_exported():
    lib = P.load()
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(P.EXPORTS) == names


def _c_layout():
    """sizeof/offsetof of the header structs, from a C program built with gcc."""
    prog = r"""
#include <stdio.h>
#include <stddef.h>
#include "powspec_hip.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m))
int main(void) {
  printf("psh_data %zu\npsh_cata %zu\npsh_conf %zu\npsh_pk %zu\n", sizeof(psh_data),
         sizeof(psh_cata), sizeof(psh_conf), sizeof(psh_pk));
  F(psh_cata, data); F(psh_cata, norm);
  F(psh_conf, gsize); F(psh_conf, poles); F(psh_conf, bsize); F(psh_conf, los);
  F(psh_conf, kmin); F(psh_conf, kbin); F(psh_conf, isauto); F(psh_conf, verbose);
  F(psh_pk, k); F(psh_pk, cnt); F(psh_pk, pl); F(psh_pk, xpl);
  printf("codes %d %d %d %d %d %d %d %d %d\n", PSH_ERR_MEMORY, PSH_ERR_ARG, PSH_ERR_FILE,
         PSH_ERR_CATA, PSH_ERR_CNVT, PSH_ERR_MESH, PSH_ERR_PK, PSH_ERR_DEVICE,
         PSH_ERR_UNSUPPORTED);
  printf("psh_cnvt %zu\n", sizeof(psh_cnvt));
  printf("psh_box_spec %zu\npsh_sky_spec %zu\n", sizeof(psh_box_spec), sizeof(psh_sky_spec));
  printf("psh_tune %zu\n", sizeof(psh_tune));
  printf("psh_cat_header %zu\npsh_ingest_result %zu\n", sizeof(psh_cat_header),
         sizeof(psh_ingest_result));
  F(psh_cat_header, n); F(psh_cat_header, dtype); F(psh_cat_header, cols);
  F(psh_ingest_result, sumw2n); F(psh_ingest_result, seconds);
  F(psh_tune, ms); F(psh_tune, ncand);
  F(psh_box_spec, seed); F(psh_sky_spec, nbar0); F(psh_sky_spec, seed);
  F(psh_cnvt, zsp); F(psh_cnvt, nsp); F(psh_cnvt, rcnvt);
  printf("psh_mesh_meta %zu\n", sizeof(psh_mesh_meta));
  F(psh_mesh_meta, bmin); F(psh_mesh_meta, intlace);
  return 0;
}
"""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(c, "w").write(prog)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), c, "-o", exe],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    vals = {}
    for ln in out.splitlines():
        k, *v = ln.split()
        vals[k] = [int(x) for x in v]
    return vals


def test_struct_layout_matches_header():
    v = _c_layout()
    assert v["psh_data"][0] == 32
    assert v["psh_cata"][0] == C.sizeof(P.Cata)
    assert v["psh_conf"][0] == C.sizeof(P.Conf)
    assert v["psh_pk"][0] == C.sizeof(P.Pk)
    for tname, cls in (("psh_cata", P.Cata), ("psh_conf", P.Conf), ("psh_pk", P.Pk),
                       ("psh_mesh_meta", P.MeshMeta), ("psh_cnvt", P.Cnvt),
                       ("psh_box_spec", P.BoxSpec), ("psh_sky_spec", P.SkySpec),
                       ("psh_tune", P.Tune)):
        for key, (off,) in ((k, o) for k, o in v.items() if k.startswith(tname + ".")):
            field = key.split(".", 1)[1]
            assert getattr(cls, field).offset == off, key
    assert v["codes"] == [P.ERR_MEMORY, P.ERR_ARG, P.ERR_FILE, P.ERR_CATA, P.ERR_CNVT,
                          P.ERR_MESH, P.ERR_PK, P.ERR_DEVICE, P.ERR_UNSUPPORTED]
    assert v["psh_cnvt"][0] == C.sizeof(P.Cnvt)
    assert v["psh_box_spec"][0] == C.sizeof(P.BoxSpec)
    assert v["psh_sky_spec"][0] == C.sizeof(P.SkySpec)
    assert v["psh_tune"][0] == C.sizeof(P.Tune)
    assert v["psh_cat_header"][0] == C.sizeof(P.CatHeader) == 64
    assert v["psh_ingest_result"][0] == C.sizeof(P.IngestResult)
    for f in ("n", "dtype", "cols"):
        assert v[f"psh_cat_header.{f}"][0] == getattr(P.CatHeader, f).offset
    for f in ("sumw2n", "seconds"):
        assert v[f"psh_ingest_result.{f}"][0] == getattr(P.IngestResult, f).offset
    assert v["psh_mesh_meta"][0] == C.sizeof(P.MeshMeta)


def test_conf_validation_needs_no_device():
    """Invalid configurations are refused before any device work."""
    lib = P.load()
    h = C.c_void_p()
    bad = P.make_conf(gsize=100, bsize=1000.0, kbin=0.01)    # not a power of two
    rc = lib.psh_ctx_create(C.byref(bad), 1, C.byref(h))
    assert rc in (P.ERR_UNSUPPORTED, P.ERR_DEVICE)
    assert not h.value


@pytest.mark.skipif(P.device_count() > 0, reason="a GPU is visible")
def test_no_gpu_fails_loudly():
    """Without a GPU the library reports PSH_ERR_DEVICE: there is no CPU path."""
    assert P.device_count() == 0
    with pytest.raises(P.PshError) as e:
        P.Context(P.make_conf(gsize=16, bsize=100.0, kbin=0.05))
    assert e.value.code == P.ERR_DEVICE
    with pytest.raises(P.PshError):
        P.DeviceArray.from_numpy(np.zeros((4, 4)))
