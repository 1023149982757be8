"""Python bindings (ctypes) of the gfx950 power-spectrum library.

The product is the C-ABI library ``libpowspec_hip.so`` (include/powspec_hip.h);
this module only marshals arguments for tests, benchmarks and Python hosts.
It never falls back to a CPU implementation: when the library is missing or
cannot load, ``load()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpowspec_hip.so")

NGP, CIC, TSC, PCS = 0, 1, 2, 3
MAX_ELL = 6
OK = 0
ERR_MEMORY, ERR_ARG, ERR_FILE, ERR_CATA, ERR_MESH, ERR_PK = -1, -2, -3, -11, -13, -14
ERR_CNVT = -12
ERR_DEVICE, ERR_UNSUPPORTED = -20, -21
FAULT_ALLOC, FAULT_GRAPH = 1, 2


class Conf(C.Structure):
    """psh_conf (include/powspec_hip.h)."""

    _fields_ = [
        ("ndata", C.c_int), ("issim", C.c_int), ("gsize", C.c_int),
        ("assign", C.c_int), ("intlace", C.c_int), ("npole", C.c_int),
        ("poles", C.c_int * (MAX_ELL + 1)), ("has_bsize", C.c_int),
        ("bsize", C.c_double * 3), ("bpad", C.c_double * 3),
        ("los", C.c_double * 3), ("logscale", C.c_int),
        ("kmin", C.c_double), ("kmax", C.c_double), ("kbin", C.c_double),
        ("isauto", C.c_int * 2), ("iscross", C.c_int), ("verbose", C.c_int),
    ]


class Pk(C.Structure):
    """psh_pk (include/powspec_hip.h)."""

    _fields_ = [
        ("nbin", C.c_int), ("nl", C.c_int), ("poles", C.c_int * (MAX_ELL + 1)),
        ("k", C.c_void_p), ("kedge", C.c_void_p), ("km", C.c_void_p),
        ("cnt", C.c_void_p), ("lcnt", C.c_void_p), ("pl", C.c_void_p * 2),
        ("xpl", C.c_void_p),
    ]


class Cata(C.Structure):
    """psh_cata: same layout as the reference's CATA."""

    _fields_ = [
        ("num", C.c_int), ("data", C.c_void_p), ("rand", C.c_void_p),
        ("ndata", C.c_void_p), ("nrand", C.c_void_p), ("wdata", C.c_void_p),
        ("wrand", C.c_void_p), ("alpha", C.c_void_p), ("shot", C.c_void_p),
        ("norm", C.c_void_p),
    ]


EXPORTS = [
    "psh_genr_mesh", "psh_powspec", "psh_mesh_destroy", "psh_powspec_destroy",
    "psh_mesh_info", "psh_ctx_create", "psh_ctx_destroy", "psh_ctx_set_box",
    "psh_ctx_nbin", "psh_ctx_run_host", "psh_ctx_define_box", "psh_ctx_run_device",
    "psh_ctx_run_batch", "psh_ctx_timings", "psh_ctx_set_profiling", "psh_last_error",
    "psh_device_count", "psh_device_malloc", "psh_device_free", "psh_memcpy_h2d",
    "psh_memcpy_d2h", "psh_memcpy_d2d", "psh_device_synchronize", "psh_device_arch", "psh_write_pk",
    "psh_save_res", "psh_cnvt_coord", "psh_cnvt_device", "psh_gen_philox",
    "psh_gen_uniform_box", "psh_gen_clustered_box", "psh_gen_sky", "psh_ctx_set_tiles",
    "psh_ctx_get_tiles",
    "psh_ctx_get_box",
    "psh_ctx_load_tuning", "psh_autotune", "psh_ctx_run_survey_device", "psh_mlp_param_count",
    "psh_mlp_forward", "psh_mlp_grad", "psh_mlp_train", "psh_ingest_create",
    "psh_ingest_destroy", "psh_cat_write", "psh_ingest_file", "psh_ingest_host",
    "psh_ingest_wait", "psh_ingest_error", "psh_ingest_ascii", "psh_parse_double",
    "psh_ingest_ascii_expr",
    "psh_fault_inject", "psh_run_device_static", "psh_ctx_set_mu_cells", "psh_ctx_cells",
    "psh_ctx_alias_rows", "psh_ctx_set_alias_mlp", "psh_ctx_set_fft_schedule",
    "psh_ctx_get_fft_schedule", "psh_autotune_fft", "psh_fft_plan", "psh_fused_tile",
]


class Tune(C.Structure):
    """psh_tune: result of psh_autotune."""

    _fields_ = [("tile", C.c_int * 3), ("ms", C.c_double), ("ncand", C.c_int)]


class TuneFft(C.Structure):
    """psh_tune_fft: result of psh_autotune_fft."""

    _fields_ = [("ysched", C.c_int), ("ms", C.c_double), ("ncand", C.c_int),
                ("ms_default", C.c_double)]


def fused_tile(Ng, intlace):
    """The deposit tile (whole z rows) that fuses the z transform into the
    deposit for a box of Ng cells, or None (host-only)."""
    t = (C.c_int * 3)()
    return tuple(t) if load().psh_fused_tile(int(Ng), int(bool(intlace)), t) else None


def fft_plan(Ng, plan=0):
    """Radices of the passes of a length-Ng transform under radix plan 0, 1
    or 2 (host-only; [] when Ng is not a supported grid size)."""
    L = load()
    r = (C.c_int * 16)()
    n = L.psh_fft_plan(int(Ng), int(plan), r, 16)
    return [r[i] for i in range(min(n, 16))]


def autotune_fft(conf, nmock=1, cache_path=None):
    """Time the y-pass schedules for conf's grid size on nmock meshes;
    returns (schedule, ms, candidates, ms of schedule 0) and appends the
    winner to cache_path."""
    t = TuneFft()
    L = load()
    rc = L.psh_autotune_fft(C.byref(conf), int(nmock),
                            cache_path.encode() if cache_path else None, C.byref(t))
    if rc:
        raise PshError(rc, L.psh_last_error(None).decode())
    return t.ysched, t.ms, t.ncand, t.ms_default


def autotune(conf, npart=0, nmock=1, cache_path=None, data=None, offsets=None):
    """Time candidate deposit tiles for conf's grid/scheme/interlacing on the
    device catalogues `data` (DeviceArray) split by `offsets`, or on a
    generated uniform box of npart per mock; returns (tile, ms, candidates)
    and appends the winner to cache_path."""
    t = Tune()
    off = None
    if data is not None:
        off = (C.c_size_t * (nmock + 1))(*[int(o) for o in offsets])
    rc = load().psh_autotune(C.byref(conf), data.ptr if data is not None else None, off,
                             C.c_size_t(int(npart)), int(nmock),
                             cache_path.encode() if cache_path else None, C.byref(t))
    if rc:
        raise PshError(rc, "psh_autotune failed")
    return tuple(t.tile), t.ms, t.ncand



class Cnvt(C.Structure):
    """psh_cnvt: the CONF fields of the coordinate conversion."""

    _fields_ = [("omega_m", C.c_double), ("omega_l", C.c_double), ("omega_k", C.c_double),
                ("eos_w", C.c_double), ("ecdst", C.c_double), ("zsp", C.c_void_p),
                ("dsp", C.c_void_p), ("nsp", C.c_size_t), ("dcnvt", C.c_int * 2),
                ("rcnvt", C.c_int * 2)]


def make_cnvt(omega_m=0.31, omega_l=None, eos_w=-1.0, ecdst=1e-8, samples=None,
              dcnvt=(True, True), rcnvt=(True, True)):
    """psh_cnvt as load_conf sets it (omega_k = 1 - Om - OL, or 0 when
    OMEGA_LAMBDA is not given, src/load_conf.c:758-767).  samples: optional
    (z, d) arrays of Z_CMVDST_CONV; kept alive on the returned object."""
    c = Cnvt()
    c.omega_m = omega_m
    if omega_l is None:
        c.omega_l, c.omega_k = 1 - omega_m, 0.0
    else:
        c.omega_l, c.omega_k = omega_l, 1 - omega_m - omega_l
    c.eos_w, c.ecdst = eos_w, ecdst
    if samples is not None:
        z = np.ascontiguousarray(samples[0], np.float64)
        d = np.ascontiguousarray(samples[1], np.float64)
        c._keep = (z, d)
        c.zsp, c.dsp, c.nsp = z.ctypes.data, d.ctypes.data, len(z)
    c.dcnvt[:] = [int(v) for v in dcnvt]
    c.rcnvt[:] = [int(v) for v in rcnvt]
    return c


def cnvt_device(cv: Cnvt, arrays):
    """Convert device arrays (DeviceArray, record count) in place; returns the
    integration order (0 for interpolation)."""
    n = len(arrays)
    ptrs = (C.c_void_p * max(n, 1))(*[a.ptr for a, _ in arrays])
    cnt = (C.c_size_t * max(n, 1))(*[int(k) for _, k in arrays])
    order = C.c_int()
    rc = load().psh_cnvt_device(C.byref(cv), ptrs, cnt, n, C.byref(order))
    if rc:
        raise PshError(rc, "coordinate conversion failed")
    return order.value


class MeshMeta(C.Structure):
    """psh_mesh_meta: mesh metadata of the output header."""

    _fields_ = [("Ng", C.c_int), ("bsize", C.c_double * 3), ("bmin", C.c_double * 3),
                ("assign", C.c_int), ("intlace", C.c_int)]

_lib = None


def load():
    """Load libpowspec_hip.so (raises when it is missing: no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp, sz, i, d = C.c_void_p, C.c_size_t, C.c_int, C.c_double
    L.psh_ctx_create.argtypes = [C.POINTER(Conf), i, C.POINTER(vp)]
    L.psh_ctx_destroy.argtypes = [vp]
    L.psh_ctx_set_box.argtypes = [vp, C.POINTER(d), C.POINTER(d)]
    L.psh_ctx_nbin.argtypes = [vp]
    L.psh_ctx_get_tiles.argtypes = [vp, C.POINTER(C.c_int)]
    L.psh_ctx_get_box.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.psh_ctx_run_host.argtypes = [vp, C.POINTER(Cata), C.POINTER(Pk)]
    L.psh_ctx_define_box.argtypes = [vp, C.POINTER(Cata)]
    L.psh_ctx_run_device.argtypes = [vp, vp, sz, d, C.POINTER(Pk)]
    L.psh_ctx_run_batch.argtypes = [vp, vp, C.POINTER(sz), C.POINTER(d), i, i, C.POINTER(Pk)]
    L.psh_ctx_timings.argtypes = [vp, C.POINTER(d)]
    L.psh_ctx_set_profiling.argtypes = [vp, i]
    L.psh_last_error.argtypes = [vp]
    L.psh_last_error.restype = C.c_char_p
    L.psh_device_malloc.argtypes = [C.POINTER(vp), sz]
    L.psh_device_free.argtypes = [vp]
    L.psh_memcpy_h2d.argtypes = [vp, vp, sz]
    L.psh_memcpy_d2h.argtypes = [vp, vp, sz]
    L.psh_memcpy_d2d.argtypes = [vp, vp, sz]
    L.psh_device_arch.argtypes = [C.c_char_p, i]
    L.psh_genr_mesh.argtypes = [C.POINTER(Conf), C.POINTER(Cata)]
    L.psh_genr_mesh.restype = vp
    L.psh_powspec.argtypes = [C.POINTER(Conf), C.POINTER(Cata), vp]
    L.psh_powspec.restype = C.POINTER(Pk)
    L.psh_mesh_destroy.argtypes = [vp]
    L.psh_powspec_destroy.argtypes = [C.POINTER(Pk)]
    L.psh_mesh_info.argtypes = [vp, C.POINTER(i), C.POINTER(d), C.POINTER(d),
                                C.POINTER(i), C.POINTER(i)]
    L.psh_ingest_create.argtypes = [sz, C.POINTER(vp)]
    L.psh_ingest_destroy.argtypes = [vp]
    L.psh_cat_write.argtypes = [C.c_char_p, C.c_uint64, i, C.c_uint, C.POINTER(vp)]
    L.psh_ingest_file.argtypes = [vp, C.c_char_p, i, vp, sz]
    L.psh_ingest_host.argtypes = [vp, C.c_uint64, i, C.c_uint, C.POINTER(vp), i, vp, sz]
    L.psh_ingest_wait.argtypes = [vp, vp]
    L.psh_ingest_error.argtypes = [vp]
    L.psh_ingest_error.restype = C.c_char_p
    L.psh_ingest_ascii.argtypes = [vp, C.c_char_p, C.POINTER(AsciiFmt), i, vp, sz]
    L.psh_ingest_ascii_expr.argtypes = [vp, C.c_char_p, C.POINTER(AsciiFmt),
                                        C.POINTER(AsciiExpr), i, vp, sz]
    L.psh_parse_double.argtypes = [C.c_char_p, sz, C.POINTER(d)]
    L.psh_fault_inject.argtypes = [i, C.c_long]
    L.psh_run_device_static.argtypes = [C.POINTER(Conf), vp, sz, d, C.POINTER(Pk)]
    L.psh_ctx_set_mu_cells.argtypes = [vp, i]
    L.psh_ctx_cells.argtypes = [vp, vp]
    L.psh_ctx_alias_rows.argtypes = [vp, vp]
    L.psh_ctx_set_alias_mlp.argtypes = [vp, vp, i]
    L.psh_ctx_set_fft_schedule.argtypes = [vp, i]
    L.psh_ctx_get_fft_schedule.argtypes = [vp, C.POINTER(i)]
    L.psh_autotune_fft.argtypes = [C.POINTER(Conf), i, C.c_char_p, C.POINTER(TuneFft)]
    L.psh_fft_plan.argtypes = [i, i, C.POINTER(C.c_int), i]
    L.psh_fused_tile.argtypes = [i, i, C.POINTER(C.c_int)]
    _lib = L
    return L


def fault_inject(what: int, value: int) -> None:
    """psh_fault_inject: FAULT_ALLOC n > 0 fails the n-th device allocation
    of a context from now on, n < 0 every one, 0 none; FAULT_GRAPH != 0 fails
    batch-graph instantiation."""
    rc = load().psh_fault_inject(int(what), int(value))
    if rc:
        raise PshError(rc, "psh_fault_inject failed")


def memcpy_d2d(dst: "DeviceArray", dst_off: int, src: "DeviceArray", nbytes: int,
               src_off: int = 0) -> None:
    """Copy nbytes between device allocations (byte offsets)."""
    if dst_off + nbytes > dst.nbytes or src_off + nbytes > src.nbytes:
        raise ValueError("device copy out of range")
    rc = load().psh_memcpy_d2d(C.c_void_p(dst.ptr.value + dst_off),
                               C.c_void_p(src.ptr.value + src_off), C.c_size_t(nbytes))
    if rc:
        raise PshError(rc, "device to device copy failed")


def device_count() -> int:
    return int(load().psh_device_count())


def device_arch() -> str:
    buf = C.create_string_buffer(64)
    if load().psh_device_arch(buf, 64) != 0:
        return ""
    return buf.value.decode()


class PshError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def make_conf(**kw) -> Conf:
    """Conf from keyword arguments (names as in psh_conf)."""
    c = Conf()
    poles = kw.pop("poles", [0, 2, 4])
    c.npole = len(poles)
    for n, p in enumerate(poles):
        c.poles[n] = p
    c.ndata = kw.pop("ndata", 1)
    c.issim = int(kw.pop("issim", True))
    c.gsize = kw.pop("gsize")
    c.assign = kw.pop("assign", CIC)
    c.intlace = int(kw.pop("intlace", False))
    bsize = kw.pop("bsize", None)
    c.has_bsize = int(bsize is not None)
    if bsize is not None:
        c.bsize[:] = list(np.broadcast_to(np.asarray(bsize, float), (3,)))
    c.bpad[:] = list(np.broadcast_to(np.asarray(kw.pop("bpad", 0.02), float), (3,)))
    c.los[:] = list(np.broadcast_to(np.asarray(kw.pop("los", [0, 0, 1]), float), (3,)))
    c.logscale = int(kw.pop("logscale", False))
    c.kmin = kw.pop("kmin", 0.0)
    c.kmax = kw.pop("kmax", -1.0)
    c.kbin = kw.pop("kbin")
    ia = kw.pop("isauto", [True, False])
    c.isauto[0], c.isauto[1] = int(ia[0]), int(ia[1])
    c.iscross = int(kw.pop("iscross", False))
    c.verbose = int(kw.pop("verbose", False))
    if kw:
        raise TypeError(f"unknown conf keys {sorted(kw)}")
    return c


class DeviceArray:
    """A raw device allocation owned by the library's allocator."""

    def __init__(self, nbytes: int):
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        rc = load().psh_device_malloc(C.byref(self.ptr), max(1, self.nbytes))
        if rc != 0:
            raise PshError(rc, f"device allocation of {nbytes} bytes failed")

    @classmethod
    def empty(cls, shape, dtype=np.float64) -> "DeviceArray":
        return cls(int(np.prod(shape)) * np.dtype(dtype).itemsize)

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if a.nbytes and load().psh_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes) != 0:
            raise RuntimeError("host to device copy failed")
        return d

    def to_numpy(self, shape, dtype=np.float64) -> np.ndarray:
        out = np.empty(shape, dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("requested more bytes than the allocation holds")
        rc = load().psh_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes)
        if rc:
            raise PshError(rc, "device to host copy failed")
        return out

    def free(self):
        if self.ptr:
            load().psh_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _PkArrays:
    def __init__(self, nbin, nl):
        self.k = np.zeros(nbin)
        self.kedge = np.zeros(nbin + 1)
        self.km = np.zeros(nbin)
        self.cnt = np.zeros(nbin, np.uint64)
        self.lcnt = np.zeros(nl * nbin)
        self.pl = [np.zeros(nl * nbin), np.zeros(nl * nbin)]
        self.xpl = np.zeros(nl * nbin)
        self.pk = Pk()
        self.pk.k = self.k.ctypes.data
        self.pk.kedge = self.kedge.ctypes.data
        self.pk.km = self.km.ctypes.data
        self.pk.cnt = self.cnt.ctypes.data
        self.pk.lcnt = self.lcnt.ctypes.data
        self.pk.pl[0] = self.pl[0].ctypes.data
        self.pk.pl[1] = self.pl[1].ctypes.data
        self.pk.xpl = self.xpl.ctypes.data
        self.nbin, self.nl = nbin, nl

    def result(self):
        nb, nl = self.nbin, self.nl
        return {"nbin": nb, "k": self.k.copy(), "kedge": self.kedge.copy(),
                "km": self.km.copy(), "cnt": self.cnt.copy(),
                "lcnt": self.lcnt.reshape(nl, nb).copy(),
                "pl": [p.reshape(nl, nb).copy() for p in self.pl],
                "xpl": self.xpl.reshape(nl, nb).copy()}


class Context:
    """A psh_ctx: one configuration on the current device."""

    def __init__(self, conf: Conf, nmock_max: int = 1):
        self.L = load()
        self.conf = conf
        self.h = C.c_void_p()
        rc = self.L.psh_ctx_create(C.byref(conf), nmock_max, C.byref(self.h))
        if rc != 0:
            raise PshError(rc, self.L.psh_last_error(None).decode())

    def _check(self, rc):
        if rc != 0:
            raise PshError(rc, self.L.psh_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.L.psh_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_profiling(self, on=True):
        self._check(self.L.psh_ctx_set_profiling(self.h, int(on)))

    def timings(self):
        t = (C.c_double * 8)()
        self._check(self.L.psh_ctx_timings(self.h, t))
        return list(t)

    def set_tiles(self, tile):
        self._check(self.L.psh_ctx_set_tiles(self.h, (C.c_int * 3)(*tile)))

    def set_fft_schedule(self, ysched):
        self._check(self.L.psh_ctx_set_fft_schedule(self.h, int(ysched)))

    def fft_schedule(self):
        v = C.c_int()
        self._check(self.L.psh_ctx_get_fft_schedule(self.h, C.byref(v)))
        return v.value

    def tiles(self):
        """The deposit tiles in use (clipped to the grid size)."""
        t = (C.c_int * 3)()
        self._check(self.L.psh_ctx_get_tiles(self.h, t))
        return tuple(t)

    def box(self):
        """(side lengths, lower corner) of the grid's box."""
        b, m = (C.c_double * 3)(), (C.c_double * 3)()
        self._check(self.L.psh_ctx_get_box(self.h, b, m))
        return tuple(b), tuple(m)

    def set_mu_cells(self, nmu):
        self._check(self.L.psh_ctx_set_mu_cells(self.h, int(nmu)))
        self.nmu = int(nmu)

    def cells(self):
        """Raw (pole, bin, |mu| cell) sums of the last box run."""
        nmu = getattr(self, "nmu", 1)
        out = np.zeros((self.conf.npole, self.nbin(), nmu))
        self._check(self.L.psh_ctx_cells(self.h, out.ctypes.data))
        return out

    def alias_rows(self):
        """The learned correction's input rows of the last box run."""
        nmu = getattr(self, "nmu", 1)
        X = np.zeros((self.nbin() * nmu, 32), np.float32)
        self._check(self.L.psh_ctx_alias_rows(self.h, X.ctypes.data))
        return X

    def set_alias_mlp(self, params):
        """Networks (one [NPAR] or an ensemble [n, NPAR]) of the learned
        correction; None clears it."""
        if params is None:
            self._check(self.L.psh_ctx_set_alias_mlp(self.h, None, 0))
            return
        p = np.ascontiguousarray(np.atleast_2d(params), np.float32)
        self._check(self.L.psh_ctx_set_alias_mlp(self.h, p.ctypes.data, len(p)))

    def load_tuning(self, cache_path):
        self._check(self.L.psh_ctx_load_tuning(self.h, str(cache_path).encode()))

    def set_box(self, bsize, bmin=(0.0, 0.0, 0.0)):
        b = (C.c_double * 3)(*np.broadcast_to(np.asarray(bsize, float), (3,)))
        m = (C.c_double * 3)(*bmin)
        self._check(self.L.psh_ctx_set_box(self.h, b, m))

    def nbin(self):
        return int(self.L.psh_ctx_nbin(self.h))

    def run_device(self, d_data: DeviceArray, n: int, weight_sum: float):
        nb = self.nbin()
        arr = _PkArrays(nb, self.conf.npole)
        self._check(self.L.psh_ctx_run_device(self.h, d_data.ptr, n, float(weight_sum),
                                              C.byref(arr.pk)))
        return arr.result()

    def run_survey_device(self, d_data: DeviceArray, nd: int, d_rand: DeviceArray, nr: int,
                          alpha: float, shot: float, norm: float):
        nb = None
        arr = None
        # the box (and so nbin) is known only after the bounds pass: size the
        # output generously, then trim
        cap = 1 << 14
        arr = _PkArrays(cap, self.conf.npole)
        self._check(self.L.psh_ctx_run_survey_device(
            self.h, d_data.ptr, C.c_size_t(nd), d_rand.ptr, C.c_size_t(nr), C.c_double(alpha),
            C.c_double(shot), C.c_double(norm), C.byref(arr.pk)))
        nb = self.nbin()
        res = arr.result()
        out = {"nbin": nb, "k": res["k"][:nb], "kedge": res["kedge"][:nb + 1],
               "km": res["km"][:nb], "cnt": res["cnt"][:nb]}
        flat = lambda a: a.ravel()[: self.conf.npole * nb].reshape(self.conf.npole, nb)
        out["pl"] = [flat(p) for p in res["pl"]]
        out["xpl"] = flat(res["xpl"])
        return out

    def run_batch(self, d_data: DeviceArray, offsets, weight_sums, graph=False):
        nm = len(offsets) - 1
        nb = self.nbin()
        arrs = [_PkArrays(nb, self.conf.npole) for _ in range(nm)]
        pks = (Pk * nm)(*[a.pk for a in arrs])
        off = (C.c_size_t * (nm + 1))(*[int(o) for o in offsets])
        ws = (C.c_double * nm)(*[float(w) for w in weight_sums])
        self._check(self.L.psh_ctx_run_batch(self.h, d_data.ptr, off, ws, nm, int(graph), pks))
        out = []
        for m, a in enumerate(arrs):
            # the library wrote through the copies in pks; re-read the arrays
            out.append(a.result())
        return out

    def run_host(self, data, rand=None, sums=None):
        """data/rand: lists of (n,4) float64 arrays; sums: per catalogue dict
        with wdata, wrand, alpha, shot, norm (surveys)."""
        num = len(data)
        keep = []
        dptrs = (C.c_void_p * 2)()
        rptrs = (C.c_void_p * 2)()
        nd = (C.c_size_t * 2)()
        nr = (C.c_size_t * 2)()
        vals = {k: (C.c_double * 2)() for k in ("wdata", "wrand", "alpha", "shot", "norm")}
        for c in range(num):
            d = np.ascontiguousarray(data[c], np.float64)
            keep.append(d)
            dptrs[c] = d.ctypes.data
            nd[c] = len(d)
            vals["wdata"][c] = float(d[:, 3].sum()) if len(d) else 0.0
            if rand is not None:
                r = np.ascontiguousarray(rand[c], np.float64)
                keep.append(r)
                rptrs[c] = r.ctypes.data
                nr[c] = len(r)
            if sums is not None:
                for k in vals:
                    if k in sums[c]:
                        vals[k][c] = sums[c][k]
        cat = Cata()
        cat.num = num
        cat.data = C.cast(dptrs, C.c_void_p)
        cat.rand = C.cast(rptrs, C.c_void_p)
        cat.ndata = C.cast(nd, C.c_void_p)
        cat.nrand = C.cast(nr, C.c_void_p)
        for k, v in vals.items():
            setattr(cat, k, C.cast(v, C.c_void_p))
        self._check(self.L.psh_ctx_define_box(self.h, C.byref(cat)))
        arr = _PkArrays(self.nbin(), self.conf.npole)
        self._check(self.L.psh_ctx_run_host(self.h, C.byref(cat), C.byref(arr.pk)))
        res = arr.result()
        res["shot"] = [vals["shot"][0], vals["shot"][1]]
        res["norm"] = [vals["norm"][0], vals["norm"][1]]
        return res


def run_device_static(conf: Conf, d_data: DeviceArray, n: int, weight_sum: float, nbin: int):
    """psh_run_device_static: a box run with no device allocation, every
    intermediate buffer in the library's static device scratch; `nbin` is
    the capacity of the result arrays."""
    arr = _PkArrays(nbin, conf.npole)
    arr.pk.nbin = nbin
    L = load()
    rc = L.psh_run_device_static(C.byref(conf), d_data.ptr, C.c_size_t(n),
                                 C.c_double(weight_sum), C.byref(arr.pk))
    if rc:
        raise PshError(rc, L.psh_last_error(None).decode())
    res = arr.result()
    nb = arr.pk.nbin
    return {k: (v[:nb] if k in ("k", "km", "cnt") else v) for k, v in res.items()}


# ---- synthetic catalogues on the GPU (psh_gen_*) ---------------------------

class BoxSpec(C.Structure):
    """psh_box_spec: clustered box generator parameters."""

    _fields_ = [("ng", C.c_int), ("L", C.c_double), ("nbar", C.c_double),
                ("npart", C.c_double), ("f", C.c_double), ("A", C.c_double),
                ("nidx", C.c_double), ("kc", C.c_double), ("seed", C.c_uint64)]


class SkySpec(C.Structure):
    """psh_sky_spec: survey sky generator parameters."""

    _fields_ = [("ra", C.c_double * 2), ("dec", C.c_double * 2), ("zmu", C.c_double),
                ("zsigma", C.c_double), ("zlo", C.c_double), ("zhi", C.c_double),
                ("nbar0", C.c_double), ("p0", C.c_double), ("seed", C.c_uint64)]


def gen_philox(n, seed, stream, draw=0) -> np.ndarray:
    """(n, 4) uint32 Philox words of counters (i, stream, draw) on the GPU."""
    d = DeviceArray(16 * n)
    rc = load().psh_gen_philox(d.ptr, C.c_size_t(n), C.c_uint64(seed), C.c_uint32(stream),
                               C.c_uint32(draw))
    if rc:
        raise PshError(rc, "psh_gen_philox failed")
    return d.to_numpy((n, 4), np.uint32)


def gen_uniform_box(n, L=1000.0, seed=1) -> DeviceArray:
    d = DeviceArray(32 * n)
    Ls = (C.c_double * 3)(*np.broadcast_to(np.asarray(L, float), (3,)))
    rc = load().psh_gen_uniform_box(d.ptr, C.c_size_t(n), Ls, C.c_uint64(seed))
    if rc:
        raise PshError(rc, "psh_gen_uniform_box failed")
    return d


class _Owned(DeviceArray):
    """A device pointer allocated inside the library."""

    def __init__(self, ptr, nbytes):   # noqa: D401 - no allocation here
        self.ptr = ptr
        self.nbytes = nbytes


def gen_clustered_box(ng=512, L=1000.0, nbar=0.02, f=0.8, seed=2, A=1e4, n=-1.5, kc=0.5,
                      npart=0.0):
    """(DeviceArray, count) of the clustered box (same recipe and streams as
    synth.clustered_box)."""
    spec = BoxSpec(ng, L, nbar, npart or 0.0, f, A, n, kc, seed)
    ptr = C.c_void_p()
    cnt = C.c_size_t()
    rc = load().psh_gen_clustered_box(C.byref(spec), C.byref(ptr), C.byref(cnt))
    if rc:
        raise PshError(rc, "psh_gen_clustered_box failed")
    return _Owned(ptr, 32 * max(cnt.value, 1)), cnt.value


def gen_sky(n, seed, ra=(0.0, 60.0), dec=(-30.0, 30.0), zmu=0.55, zsigma=0.08, zlo=0.4,
            zhi=0.7, nbar0=3e-4, p0=1e4):
    """(DeviceArray of (RA, Dec, z, w_fkp), DeviceArray of n(z))."""
    spec = SkySpec((C.c_double * 2)(*ra), (C.c_double * 2)(*dec), zmu, zsigma, zlo, zhi,
                   nbar0, p0, seed)
    d, nz = DeviceArray(32 * n), DeviceArray(8 * n)
    rc = load().psh_gen_sky(C.byref(spec), d.ptr, nz.ptr, C.c_size_t(n))
    if rc:
        raise PshError(rc, "psh_gen_sky failed")
    return d, nz


# ---- catalogue ingest (psh_ingest_*) ---------------------------------------

COL_X, COL_Y, COL_Z, COL_WCOMP, COL_WFKP, COL_NZ = 1, 2, 4, 8, 16, 32
_COL_NAMES = ("x", "y", "z", "w_comp", "w_fkp", "nz")
CAT_MAGIC = b"PSHCAT01"


class CatHeader(C.Structure):
    """psh_cat_header: the 64-byte header of a binary catalogue file."""

    _fields_ = [("magic", C.c_char * 8), ("n", C.c_uint64), ("dtype", C.c_uint32),
                ("cols", C.c_uint32), ("reserved", C.c_uint8 * 40)]


class IngestResult(C.Structure):
    """psh_ingest_result."""

    _fields_ = [("n", C.c_uint64), ("sumw", C.c_double), ("sumw2", C.c_double),
                ("sumw2n", C.c_double), ("bytes", C.c_double), ("seconds_read", C.c_double),
                ("seconds", C.c_double)]


def _columns(cols: dict, dtype):
    """Contiguous arrays of the given dtype for the named columns; (mask,
    pointer array, keep-alive list)."""
    mask, ptrs, keep = 0, (C.c_void_p * 6)(), []
    n = None
    for c, name in enumerate(_COL_NAMES):
        if cols.get(name) is None:
            continue
        a = np.ascontiguousarray(cols[name], dtype=dtype)
        if n is None:
            n = len(a)
        elif len(a) != n:
            raise ValueError("columns of different lengths")
        mask |= 1 << c
        ptrs[c] = a.ctypes.data
        keep.append(a)
    return mask, ptrs, keep, n or 0


def cat_write(path, cols: dict, dtype=np.float64):
    """Write a binary catalogue from columns {"x": ..., "y": ..., "z": ...,
    optional "w_comp", "w_fkp", "nz"}."""
    mask, ptrs, keep, n = _columns(cols, dtype)
    rc = load().psh_cat_write(str(path).encode(), C.c_uint64(n), np.dtype(dtype).itemsize,
                              C.c_uint(mask), ptrs)
    if rc:
        raise PshError(rc, f"psh_cat_write failed for {path}")


class AsciiFmt(C.Structure):
    """psh_ascii_fmt: DATA_SKIP, DATA_COMMENT, DATA_FORMATTER and the columns
    of DATA_POSITION / DATA_WT_COMP / DATA_WT_FKP / DATA_NZ."""

    _fields_ = [("skip", C.c_long), ("comment", C.c_char), ("fmtr", C.c_char_p),
                ("pos", C.c_int * 3), ("wcomp", C.c_int), ("wfkp", C.c_int), ("nz", C.c_int)]


class AsciiExpr(C.Structure):
    """psh_ascii_expr: the libast expressions of DATA_POSITION, DATA_WT_COMP,
    DATA_WT_FKP, DATA_NZ and DATA_SELECTION."""

    _fields_ = [("pos", C.c_char_p * 3), ("wcomp", C.c_char_p), ("wfkp", C.c_char_p),
                ("nz", C.c_char_p), ("sel", C.c_char_p)]


def _column(v) -> int:
    """Column number from an int or a plain libast reference "$n" / "${n}"."""
    if v is None:
        return 0
    if isinstance(v, (int, np.integer)):
        return int(v)
    t = str(v).strip()
    if t.startswith("${") and t.endswith("}"):
        t = t[2:-1]
    elif t.startswith("$"):
        t = t[1:]
    if not t.isdigit():
        raise ValueError(f"start_ascii takes plain column references (start_ascii_expr "
                         f"takes expressions), got {v!r}")
    return int(t)


def parse_double(token: str) -> float:
    """The ASCII ingest's decimal conversion (host build of the device code)."""
    b = token.encode()
    out = C.c_double()
    rc = load().psh_parse_double(b, len(b), C.byref(out))
    if rc:
        raise PshError(rc, f"not a supported decimal number: {token!r}")
    return out.value


class Ingest:
    """psh_ingest: pinned double-buffered catalogue upload on its own stream
    and host thread.  start_file / start_host return at once; wait() gives the
    psh_ingest_result as a dict."""

    def __init__(self, chunk_objects=0):
        self.L = load()
        self.h = C.c_void_p()
        rc = self.L.psh_ingest_create(C.c_size_t(chunk_objects), C.byref(self.h))
        if rc:
            raise PshError(rc, "psh_ingest_create failed")
        self._keep = None

    def _check(self, rc):
        if rc:
            raise PshError(rc, self.L.psh_ingest_error(self.h).decode())

    def start_file(self, path, out: "DeviceArray", cap, sim=True):
        self._check(self.L.psh_ingest_file(self.h, str(path).encode(), int(bool(sim)),
                                           out.ptr, C.c_size_t(cap)))

    def start_host(self, cols: dict, out: "DeviceArray", cap, sim=True, dtype=np.float64):
        mask, ptrs, keep, n = _columns(cols, dtype)
        self._keep = (keep, ptrs)          # alive until wait()
        self._check(self.L.psh_ingest_host(self.h, C.c_uint64(n), np.dtype(dtype).itemsize,
                                           C.c_uint(mask), ptrs, int(bool(sim)), out.ptr,
                                           C.c_size_t(cap)))

    def start_ascii(self, path, out: "DeviceArray", cap, fmtr="%lf %lf %lf", pos=(1, 2, 3),
                    wcomp=None, wfkp=None, nz=None, skip=0, comment="#", sim=True):
        f = AsciiFmt()
        f.skip = int(skip)
        f.comment = (comment or "\0").encode()[:1]
        self._keep = fmtr.encode()
        f.fmtr = self._keep
        for a in range(3):
            f.pos[a] = _column(pos[a])
        f.wcomp, f.wfkp, f.nz = _column(wcomp), _column(wfkp), _column(nz)
        self._check(self.L.psh_ingest_ascii(self.h, str(path).encode(), C.byref(f),
                                            int(bool(sim)), out.ptr, C.c_size_t(cap)))

    def start_ascii_expr(self, path, out: "DeviceArray", cap, fmtr="%lf %lf %lf",
                         pos=("$1", "$2", "$3"), wcomp=None, wfkp=None, nz=None, sel=None,
                         skip=0, comment="#", sim=True):
        """ASCII catalogue with libast expressions for the positions, weights
        and n(z) and an optional selection (psh_ingest_ascii_expr)."""
        f = AsciiFmt()
        f.skip = int(skip)
        f.comment = (comment or "\0").encode()[:1]
        enc = [fmtr.encode()] + [None if e is None else str(e).encode()
                                 for e in (*pos, wcomp, wfkp, nz, sel)]
        self._keep = enc
        f.fmtr = enc[0]
        e = AsciiExpr()
        for a in range(3):
            e.pos[a] = enc[1 + a]
        e.wcomp, e.wfkp, e.nz, e.sel = enc[4], enc[5], enc[6], enc[7]
        self._check(self.L.psh_ingest_ascii_expr(self.h, str(path).encode(), C.byref(f),
                                                 C.byref(e), int(bool(sim)), out.ptr,
                                                 C.c_size_t(cap)))

    def wait(self):
        r = IngestResult()
        rc = self.L.psh_ingest_wait(self.h, C.byref(r))
        self._keep = None
        self._check(rc)
        return {f: getattr(r, f) for f, _ in IngestResult._fields_}

    def close(self):
        if self.h:
            self.L.psh_ingest_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
