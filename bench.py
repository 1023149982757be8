#!/usr/bin/env python3
"""Benchmark of the mesh -> FFT -> P_l(k) path on one MI355X.

Workload (default): the clustered periodic box at production size --
Gaussian field on a 512^3 grid with P(k) = 1e4 (k/0.1)^-1.5 exp(-(k/0.5)^2)
in L = 1000 Mpc/h, lognormal, Poisson-sampled at 0.02 (h/Mpc)^3 (~2e7
particles), Zel'dovich RSD along z with f = 0.8, generated on the GPU from
Philox streams (seed 2); PCS, N_g = 1024, interlaced, FP32 meshes and FFTs,
FP64 bin sums, l = 0,2,4, dk = 0.005.

A step is one full pass of the path over the catalogue already resident in
HBM: coordinate bounds, mass assignment of both interlaced meshes, 3-D r2c
FFTs, interlace combine, binning and normalisation (psh_ctx_run_device).

Printed: one JSON line with meshes/s (value), particles/s, the roofline of
the dominant kernel (HIP events on the library stream), batched-vs-plain
throughput of many small mocks, and the CPU oracle baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particles/s + P_l(k) meshes/s at N_g=1024, batched vs plain; % of HBM peak"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def box_conf(P):
    return dict(gsize=1024, assign=P.PCS, intlace=True, bsize=1000.0, kbin=0.005,
                poles=[0, 2, 4])


def x_live_fraction(Ng, L, kedge_max, C=16):
    """Share of the stored modes the fused x pass reads (and the y pass
    writes): the (k_y, k_z-tile) columns of C k_z values whose nearest mode
    lies before the last bin edge (same test as the library; L the side of a
    cubic box or the three sides)."""
    Ls = np.broadcast_to(np.asarray(L, float), (3,))
    M = Ng // 2
    ky = np.fft.fftfreq(Ng, 1.0 / Ng) * (2 * np.pi / Ls[1])
    q = np.arange((M + 1 + C - 1) // C)
    kz0 = q * C * (2 * np.pi / Ls[2])
    live = np.sqrt(ky[:, None] ** 2 + kz0[None, :] ** 2) < kedge_max
    cols = np.minimum(C, M + 1 - q * C)
    return float((live * cols[None, :]).sum() / (Ng * (M + 1)))


def stage_bytes(Ng, nmesh, npart, x_live=1.0):
    """Algorithmic (compulsory) HBM bytes per launch of each stage."""
    ntot = float(Ng) ** 3
    nc = float(Ng) * Ng * (Ng // 2 + 1)
    return {
        "deposit": nmesh * 4 * ntot,                    # each mesh written once
        "z_r2c": nmesh * (4 * ntot + 8 * nc),            # read real, write half-complex
        "y_fft": nmesh * 8 * nc * (1 + x_live),          # read all, write the columns the x pass reads
        "x_fft_bin": nmesh * 8 * nc * x_live,            # read the columns with binned modes
        "assign_total": npart * 32 + nmesh * 4 * ntot,   # particles once + meshes once
    }


def run_box(P, steps, warmup, fixed_tile=None, fixed_ysched=None):
    t = time.time()
    # clustered N_g=1024 box catalogue generated on the GPU (same Philox streams as powspec_amd.synth.clustered_box)
    dev, n = P.gen_clustered_box(ng=512, L=1000.0, nbar=0.02, f=0.8, seed=2)
    P.load().psh_device_synchronize()
    data = dev.to_numpy((n, 4))   # host copy for the CPU baseline and the mocks
    log(f"box catalogue: {n} particles generated in {time.time() - t:.1f} s")
    conf = box_conf(P)
    # deposit tiles tuned on this catalogue before the timed region (a planning
    # step: candidates timed on the same data; results do not depend on it)
    t = time.time()
    if fixed_tile:
        tile = tuple(fixed_tile)
        log(f"deposit tiles: {tile} (given)")
    else:
        tile, tms, ncand = P.autotune(P.make_conf(**conf), nmock=1, data=dev, offsets=[0, n])
        log(f"deposit tiles: {tile} ({tms:.2f} ms, {ncand} candidates, {time.time() - t:.1f} s)")
    if fixed_ysched is not None:
        ysched = int(fixed_ysched)
        log(f"y-pass schedule: {ysched} (given)")
    else:
        ysched, yms, yn, y0 = P.autotune_fft(P.make_conf(**conf), 1)
        log(f"y-pass schedule: {ysched} ({yms:.3f} ms per launch vs {y0:.3f} ms for "
            f"schedule 0, {yn} candidates)")
    ctx = P.Context(P.make_conf(**conf))
    ctx.set_tiles(tile)
    ctx.set_fft_schedule(ysched)
    L = P.load()
    for _ in range(warmup):
        res = ctx.run_device(dev, n, float(n))
    L.psh_device_synchronize()
    # HIP events on the library's stream bracket every stage of every timed
    # step (each step ends with the copy of its bins to the host)
    ctx.set_profiling(True)
    reps = []
    t0 = time.perf_counter()
    for _ in range(steps):
        res = ctx.run_device(dev, n, float(n))
        reps.append(ctx.timings())
    L.psh_device_synchronize()
    t1 = time.perf_counter()
    ctx.set_profiling(False)
    ms_step = (t1 - t0) / steps * 1e3
    log(f"box N_g=1024: {ms_step:.3f} ms/step over {steps} steps")
    ms = np.mean(np.array(reps), axis=0)
    stages = {"assign_total": ms[0], "zy_fft": ms[1], "x_fft_bin": ms[2], "total": ms[3],
              "deposit": ms[4], "z_r2c": ms[5], "y_fft": ms[6], "bucket": ms[7]}
    log("box stages (ms, mean over the timed steps): " +
        ", ".join(f"{k}={v:.3f}" for k, v in stages.items()))
    return data, res, ms_step, stages, n, tile, ysched


KERNELS = {"deposit": "assign_deposit", "z_r2c": "z_r2c_kernel", "y_fft": "col_kernel",
           "x_fft_bin": "x_bin_kernel"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or
    None when no summary for it is present."""
    path = os.path.join(ROOT, "profiles", "box1024_pcs_hbm_traffic.md")
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch") and kernel in ln:
                    return float(ln.split("=")[-1])
    except OSError:
        pass
    return None


def roofline(stages, Ng, nmesh, npart, x_live=1.0):
    b = stage_bytes(Ng, nmesh, npart, x_live)
    cands = {k: stages[k] for k in ("deposit", "z_r2c", "y_fft", "x_fft_bin")}
    dom = max(cands, key=lambda k: cands[k])
    per_launch = b[dom]
    # the z and y passes launch once per mesh (plain and shifted grid); the
    # deposit and the fused x pass cover both meshes in one launch
    launches = nmesh if dom in ("z_r2c", "y_fft") else 1
    t = stages[dom] / launches
    ach = per_launch / launches / (t * 1e-3) / 1e9
    table = {}
    for k in cands:
        nl = nmesh if k in ("z_r2c", "y_fft") else 1
        if stages[k] > 0:
            table[k] = {"ms": round(stages[k], 4),
                        "GBps": round(b[k] / (stages[k] * 1e-3) / 1e9, 1),
                        "launches": nl}
    return {"kernel": KERNELS[dom], "bound": "hbm", "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(KERNELS[dom]),
            "bytes_per_launch": per_launch / launches, "ms_per_launch": round(t, 4),
            "stages": table}


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain device copy of known size (torch), for scaling the HBM
    counters: FETCH_SIZE and WRITE_SIZE of this kernel should read nbytes."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                      "launches": reps}), flush=True)


def gen_mocks(P, nmock, seed=1000):
    """The batched-mock workload: nmock independent clustered boxes (the
    clustered recipe on a 256^3 generator grid, L = 1000 Mpc/h, f = 0.8),
    mock m with its own seed (seed + m) and an expected particle count drawn
    from U[2e5, 3e6], generated on the GPU and concatenated in one device
    buffer.  Returns (DeviceArray, offsets, counts)."""
    sizes = np.random.default_rng(seed).integers(200_000, 3_000_001, nmock)
    mocks, counts = [], []
    for m in range(nmock):
        dev, n = P.gen_clustered_box(ng=256, L=1000.0, f=0.8, seed=seed + m,
                                     npart=float(sizes[m]))
        mocks.append(dev)
        counts.append(n)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    allp = P.DeviceArray(32 * int(off[-1]))
    for m, dev in enumerate(mocks):
        P.memcpy_d2d(allp, 32 * int(off[m]), dev, 32 * counts[m])
        dev.free()
    P.load().psh_device_synchronize()
    return allp, off, np.array(counts)


def batched_ceiling(counts, Ng=256, nmesh=2):
    """HBM ceiling of the batched workload in meshes/s: compulsory bytes per
    mock (particles read once at 16 B, each mesh written once, one r2c
    pass reading real and writing half-complex, binning reading it) at the
    8 TB/s peak."""
    ntot = float(Ng) ** 3
    nc = float(Ng) * Ng * (Ng // 2 + 1)
    per = np.asarray(counts, float) * 16 + nmesh * (4 * ntot + (4 * ntot + 8 * nc) + 8 * nc)
    return len(counts) / (per.sum() / (HBM_PEAK_GBS * 1e9))


def run_batched(P, nmock, seed=1000, fixed_tile=None):
    """Many small clustered mocks (TSC, N_g = 256, interlaced): one batched
    launch sequence replayed as a hipGraph vs one plain run per mock.
    fixed_tile (profiling runs): deposit tiles of both instead of tuning
    them, and y-pass schedule 0."""
    t = time.time()
    dev, off, sizes = gen_mocks(P, nmock, seed)
    log(f"batched mocks: {nmock} mocks, {int(off[-1])} particles generated in "
        f"{time.time() - t:.1f} s")
    kw = dict(gsize=256, assign=P.TSC, intlace=True, bsize=1000.0, kbin=0.01, poles=[0, 2, 4])
    t = time.time()
    if fixed_tile:
        tile, ysched, ysched1 = tuple(fixed_tile), 0, 0
        log(f"batched tiles: {tile} (given), y-pass schedule 0")
    else:
        tile, tms, ncand = P.autotune(P.make_conf(**kw), nmock=nmock, data=dev, offsets=off)
        log(f"batched tiles: {tile} ({tms:.1f} ms per batch, {ncand} candidates, "
            f"{time.time() - t:.1f} s)")
        ysched = P.autotune_fft(P.make_conf(**kw), nmock)[0]
        ysched1 = P.autotune_fft(P.make_conf(**kw), 1)[0]
    ctx = P.Context(P.make_conf(**kw), nmock_max=nmock)
    ctx.set_tiles(tile)
    ctx.set_fft_schedule(ysched)
    L = P.load()
    ws = [float(s) for s in sizes]
    bres = ctx.run_batch(dev, off, ws, graph=True)    # capture + warm-up
    L.psh_device_synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.run_batch(dev, off, ws, graph=True)
    L.psh_device_synchronize()
    t_batch = (time.perf_counter() - t0) / reps
    # plain: one mock per launch sequence
    single = P.Context(P.make_conf(**kw))
    single.set_tiles(tile)
    single.set_fft_schedule(ysched1)
    ptr0 = dev.ptr.value

    class View:  # device pointer into the concatenated buffer
        def __init__(self, p):
            import ctypes
            self.ptr = ctypes.c_void_p(p)

    for m in range(min(nmock, 4)):
        single.run_device(View(ptr0 + 32 * int(off[m])), int(sizes[m]), ws[m])
    L.psh_device_synchronize()
    t0 = time.perf_counter()
    plain = []
    for m in range(nmock):
        plain.append(single.run_device(View(ptr0 + 32 * int(off[m])), int(sizes[m]), ws[m]))
    L.psh_device_synchronize()
    t_plain = time.perf_counter() - t0
    # batched == plain per mock up to the FP32/FP64 summation order:
    # |dP_l| / S_b, S_b = |P_0| + shot
    worst = 0.0
    for m in range(nmock):
        S = np.abs(plain[m]["pl"][0][0]) + 1000.0 ** 3 / sizes[m]
        for li in range(3):
            d = np.abs(bres[m]["pl"][0][li] - plain[m]["pl"][0][li]) / S
            worst = max(worst, float(np.max(d)))
    log(f"batched vs plain: worst |dP_l|/S_b over {nmock} mocks {worst:.2e}")
    ceil = batched_ceiling(sizes)
    log(f"batched: {nmock / t_batch:.1f} meshes/s ({nmock / t_batch / ceil:.3f} of the "
        f"{ceil:.0f} meshes/s HBM ceiling), plain: {nmock / t_plain:.1f} meshes/s")
    return {"workload": f"{nmock} independent clustered mocks (seeds {seed}..{seed + nmock - 1}, "
                        f"expected N_p in [2e5,3e6], total {int(off[-1])}), "
                        "TSC N_g=256 interlaced, l=0,2,4, dk=0.01",
            "batched_meshes_per_s": round(nmock / t_batch, 1),
            "plain_meshes_per_s": round(nmock / t_plain, 1),
            "batched_over_plain": round(t_plain / t_batch, 2),
            "batched_vs_plain_max_dP_over_Sb": worst,
            "tile": list(tile), "y_schedule": ysched,
            "hbm_ceiling_meshes_per_s": round(ceil, 1),
            "batched_frac_of_ceiling": round(nmock / t_batch / ceil, 4),
            "batched_particles_per_s": round(int(off[-1]) / t_batch, 1)}


def run_alias(P, steps=10, warmup=2, params_path=None, fixed_tile=None):
    """Learned alias correction: a held-out clustered box of the training
    family (P = 1e4 (k/0.1)^-1.5 exp(-(k/0.5)^2), 512^3 generator grid,
    nbar = 0.02, seed 6200), CIC, N_g = 512, l = 0,2,4, dk = 0.01; the plain
    path with the library's correction stage (8 |mu| cells, the shipped
    network ensemble) against the interlaced path.  Both timed with the
    catalogue in HBM: HIP events from the run's first to its last kernel
    (device) and wall time per call."""
    from powspec_amd import mlp
    params = mlp.load_params(params_path) if params_path else mlp.load_params()
    if params is None:
        log("alias correction: no trained networks in the package, skipped")
        return None
    dev, n = P.gen_clustered_box(ng=512, L=1000.0, nbar=0.02, f=0.8, seed=6200)
    L = P.load()
    out = {}
    for name, il in (("plain_mlp", False), ("interlaced", True)):
        conf = mlp.box_conf(P, 512, 1000.0, 0.01, P.CIC, il)
        # fixed_tile: one tile for both paths, or (plain tile, interlaced tile)
        ft = fixed_tile
        if ft and isinstance(ft[0], (list, tuple)):
            ft = ft[1] if il else ft[0]
        tile = (tuple(ft) if ft else P.autotune(conf, nmock=1, data=dev, offsets=[0, n])[0])
        ctx = P.Context(conf)
        ctx.set_tiles(tile)
        if not il:
            ctx.set_mu_cells(mlp.NMU)
            ctx.set_alias_mlp(params)
        for _ in range(warmup):
            res = ctx.run_device(dev, n, float(n))
        L.psh_device_synchronize()
        ctx.set_profiling(True)
        dev_ms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            res = ctx.run_device(dev, n, float(n))
            dev_ms.append(ctx.timings()[3])
        wall = (time.perf_counter() - t0) / steps * 1e3
        out[name] = (float(np.mean(dev_ms)), wall, res, tile)
        ctx.close()
    dev.free()
    (dp, wp, rp, tp), (di, wi, ri, ti) = out["plain_mlp"], out["interlaced"]
    kny = np.pi * 512 / 1000.0
    sel = rp["k"] <= 0.8 * kny
    shot = 1000.0 ** 3 / n
    S = ri["pl"][0][0][sel] + shot
    err = [float(np.max(np.abs(rp["pl"][0][l][sel] - ri["pl"][0][l][sel]) / S)) for l in range(3)]
    log(f"alias correction: plain+MLP {dp:.3f} ms device / {wp:.3f} ms wall, interlaced "
        f"{di:.3f} / {wi:.3f} ms (ratio {dp / di:.3f} device, {wp / wi:.3f} wall); worst "
        f"|dP_l|/S_b for k <= 0.8 k_Ny vs interlaced: {np.round(err, 5).tolist()}")
    return {"workload": f"held-out clustered box (seed 6200, {n} particles), CIC N_g=512, "
                        "l=0,2,4, dk=0.01; plain + learned correction (8 |mu| cells, "
                        f"{len(params)} networks) vs interlaced",
            "plain_mlp_ms_device": round(dp, 4), "interlaced_ms_device": round(di, 4),
            "plain_mlp_ms_wall": round(wp, 4), "interlaced_ms_wall": round(wi, 4),
            "ratio_device": round(dp / di, 4), "ratio_wall": round(wp / wi, 4),
            "max_dP_over_Sb_l024": err, "tiles": [list(tp), list(ti)]}


def run_survey(P, steps=3):
    """Survey workload: sky cap RA in [0, 60) deg, Dec in [-30, 30)
    deg (uniform sin Dec), redshift from a Gaussian n(z) (0.55, 0.08)
    truncated to [0.4, 0.7], 5e6 data + 1e8 randoms generated on the GPU,
    w_fkp = 1/(1 + n(z) 1e4), converted to comoving coordinates on the GPU
    (flat LCDM, Om = 0.31), CIC, N_g = 512, interlaced, l = 0,2,4, dk = 0.01.
    Timed: the survey estimator with both catalogues in HBM."""
    t = time.time()
    nd, nr = 5_000_000, 100_000_000
    kw = dict(ra=(0.0, 60.0), dec=(-30.0, 30.0), zmu=0.55, zsigma=0.08, zlo=0.4, zhi=0.7,
              nbar0=5e-3, p0=1e4)
    dd, dnz = P.gen_sky(nd, 4001, **kw)
    rr, rnz = P.gen_sky(nr, 4002, **kw)
    cv = P.make_cnvt(omega_m=0.31)
    order = P.cnvt_device(cv, [(dd, nd), (rr, nr)])
    P.load().psh_device_synchronize()
    t_gen = time.time() - t
    # catalogue sums (src/read_cata.c:165-183) with completeness weights 1
    wd = dd.to_numpy((nd, 4))[:, 3]
    wr = rr.to_numpy((nr, 4))[:, 3]
    nzd, nzr = dnz.to_numpy((nd,)), rnz.to_numpy((nr,))
    alpha = nd / nr
    shot = float(np.sum(wd * wd)) + alpha * alpha * float(np.sum(wr * wr))
    norm = alpha * float(np.sum(wr * wr * nzr))
    del wd, wr, nzd, nzr
    conf = P.make_conf(gsize=512, assign=P.CIC, intlace=True, issim=False, kbin=0.01,
                       poles=[0, 2, 4])
    ctx = P.Context(conf)
    res = ctx.run_survey_device(dd, nd, rr, nr, alpha, shot, norm)   # warm-up
    L = P.load()
    L.psh_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        res = ctx.run_survey_device(dd, nd, rr, nr, alpha, shot, norm)
    L.psh_device_synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    log(f"survey: {ms:.2f} ms per P_l evaluation (generation + conversion {t_gen:.1f} s, "
        f"Gauss-Legendre order {order}), nbin {res['nbin']}")
    # the l > 0 fields' fused x pass (x_bin_kernel<512, SURVEYL>): HIP events
    # around its launches in one more run.  One launch per pole (and bin
    # window) transforms all 2l + 1 fields of the pole and bins them
    # against F0: it reads each field and F0 once on the (k_y, k_z) columns
    # with binned modes
    ctx.set_profiling(True)
    ctx.run_survey_device(dd, nd, rr, nr, alpha, shot, norm)
    tm = ctx.timings()
    ctx.set_profiling(False)
    side, _ = ctx.box()
    nlm, npole_l = 5 + 9, 2
    Ng = 512
    live = x_live_fraction(Ng, side, float(res["kedge"][-1]))
    mesh_bytes = 8 * Ng * Ng * (Ng // 2 + 1) * live
    eval_bytes = (nlm + npole_l) * mesh_bytes
    dom = None
    if tm[5] > 0:
        rate = eval_bytes / (tm[5] * 1e-3) / 1e9
        dom = {"kernel": "x_bin_kernel<512, SURVEYL> (l > 0 fused x pass, one launch per pole)",
               "launches_per_evaluation": npole_l, "fields_per_evaluation": nlm,
               "ms_per_evaluation": round(tm[5], 3),
               "bytes_per_evaluation": round(eval_bytes), "live_column_fraction": round(live, 4),
               "achieved_GBps": round(rate, 1), "peak_GBps": 8000.0,
               "frac": round(rate / 8000.0, 4),
               "z_y_passes_ms_per_evaluation": round(tm[6], 3)}
        log(f"survey: l > 0 fused x passes {tm[5]:.2f} ms per evaluation ({npole_l} launches, "
            f"{nlm} fields + F0 per pole, {eval_bytes / 1e9:.3f} GB): {rate:.0f} GB/s, "
            f"{rate / 8000:.3f} of HBM peak; their z + y passes {tm[6]:.2f} ms")
    return {"workload": "survey sky cap 60x60 deg, z in [0.4,0.7] Gaussian n(z), 5e6 data + "
                        "1e8 randoms, GPU coordinate conversion, CIC N_g=512 interlaced, "
                        "l=0,2,4, dk=0.01",
            "ms_per_evaluation": round(ms, 3), "evaluations_per_s": round(1e3 / ms, 3),
            "particles_per_s": round((nd + nr) / (ms * 1e-3), 1), "nbin": res["nbin"],
            "fft_transforms": 2 + (5 + 9), "dominant_kernel": dom}


def run_dropin(P, data, reps=3):
    """One-shot drop-in call on the clustered-box catalogue as the reference's
    host would make it: psh_genr_mesh (new context, box, mode geometry, copy
    of the malloc'd catalogue to the device, which the library then frees)
    and psh_powspec (assignment, FFTs, binning), wall time per call."""
    import ctypes as C
    L = P.load()
    libc = C.CDLL("libc.so.6")
    libc.malloc.restype = C.c_void_p
    libc.malloc.argtypes = [C.c_size_t]
    L.psh_genr_mesh.restype = C.c_void_p
    L.psh_genr_mesh.argtypes = [C.c_void_p, C.c_void_p]
    L.psh_powspec.restype = C.POINTER(P.Pk)
    L.psh_powspec.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.psh_mesh_destroy.argtypes = [C.c_void_p]
    L.psh_powspec_destroy.argtypes = [C.c_void_p]
    n = len(data)
    data = np.ascontiguousarray(data)
    conf = P.make_conf(**box_conf(P))
    times = []
    for _ in range(reps):
        psz = C.sizeof(C.c_void_p)
        dptr_p, rptr_p = libc.malloc(psz), libc.malloc(psz)
        C.memset(rptr_p, 0, psz)
        p = libc.malloc(data.nbytes)
        C.memmove(p, data.ctypes.data, data.nbytes)
        (C.c_void_p * 1).from_address(dptr_p)[0] = p
        nd = (C.c_size_t * 1)(n)
        vals = {k: (C.c_double * 1)() for k in ("wdata", "wrand", "alpha", "shot", "norm")}
        vals["wdata"][0] = float(n)
        cat = P.Cata()
        cat.num = 1
        cat.data, cat.rand = dptr_p, rptr_p
        cat.ndata = C.cast(nd, C.c_void_p)
        cat.nrand = C.cast((C.c_size_t * 1)(), C.c_void_p)
        for k, v in vals.items():
            setattr(cat, k, C.cast(v, C.c_void_p))
        # the drop-in layer prints the reference's progress lines on stdout:
        # send them to stderr, so that stdout keeps the one JSON line
        sys.stdout.flush()
        libc.fflush(None)
        saved = os.dup(1)
        os.dup2(2, 1)
        try:
            t0 = time.perf_counter()
            mesh = L.psh_genr_mesh(C.byref(conf), C.byref(cat))
            t1 = time.perf_counter()
            pk = L.psh_powspec(C.byref(conf), C.byref(cat), mesh) if mesh else None
            t2 = time.perf_counter()
        finally:
            libc.fflush(None)
            os.dup2(saved, 1)
            os.close(saved)
        if not mesh or not pk:
            log("drop-in: call failed")
            return None
        L.psh_powspec_destroy(pk)
        L.psh_mesh_destroy(C.c_void_p(mesh))
        times.append((1e3 * (t1 - t0), 1e3 * (t2 - t1)))
    g = [t[0] for t in times]
    s = [t[1] for t in times]
    log(f"drop-in one-shot: genr_mesh {np.round(g, 1).tolist()} ms, powspec "
        f"{np.round(s, 1).tolist()} ms")
    return {"workload": "psh_genr_mesh + psh_powspec on the clustered-box catalogue "
                        f"({n} particles in host memory, PCS N_g=1024 interlaced), new "
                        "context per call (box, mode geometry, upload, path)",
            "genr_mesh_ms": [round(v, 2) for v in g], "powspec_ms": [round(v, 2) for v in s],
            "total_ms_median": round(float(np.median(np.add(g, s))), 2)}


def run_ingest(P, data, tile, reps=3):
    """Catalogue ingest of the clustered box catalogue (x, y, z float64 columns, unit
    weights): binary file -> HBM and host columns -> HBM through the pinned
    pipeline, and a stream of catalogues ingested while the previous one is
    processed, against ingest-then-process in sequence.  The file is written
    to the box's /tmp just before and read back from the page cache, so the
    file rate is the host-memory -> PCIe -> HBM rate, not a disk rate."""
    import tempfile
    n = len(data)
    cols = {k: np.ascontiguousarray(data[:, a]) for a, k in enumerate("xyz")}
    out = [P.DeviceArray.empty((n, 4)) for _ in range(2)]
    ing = P.Ingest()
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        path = os.path.join(d, "box.pscat")
        P.cat_write(path, cols)
        ing.start_file(path, out[0], n)
        ing.wait()
        t_file, t_host = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            ing.start_file(path, out[0], n)
            r = ing.wait()
            t_file.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            ing.start_host(cols, out[1], n)
            ing.wait()
            t_host.append(time.perf_counter() - t0)
        nbytes = r["bytes"]
        # streaming: catalogue k+1 ingested while catalogue k is processed
        ctx = P.Context(P.make_conf(**box_conf(P)))
        ctx.set_tiles(tile)
        ctx.run_device(out[0], n, float(n))
        ncat = 4
        t0 = time.perf_counter()
        for k in range(ncat):
            ing.start_file(path, out[k % 2], n)
            ing.wait()
            ctx.run_device(out[k % 2], n, float(n))
        t_seq = time.perf_counter() - t0
        t0 = time.perf_counter()
        ing.start_file(path, out[0], n)
        ing.wait()
        for k in range(ncat):
            if k + 1 < ncat:
                ing.start_file(path, out[(k + 1) % 2], n)
            ctx.run_device(out[k % 2], n, float(n))
            if k + 1 < ncat:
                ing.wait()
        t_ovl = time.perf_counter() - t0
        ctx.close()
        # ASCII (the reference's input format): 2e6 lines of x y z, %.10f
        na = min(n, 2_000_000)
        apath = os.path.join(d, "box.txt")
        np.savetxt(apath, data[:na, :3], fmt="%.10f")
        abytes = os.path.getsize(apath)
        t_ascii = []
        for _ in range(reps):
            t0 = time.perf_counter()
            ing.start_ascii(apath, out[0], n, comment="#")
            ra = ing.wait()
            t_ascii.append(time.perf_counter() - t0)
        assert ra["n"] == na
        # the same file through libast-style expressions and a selection that
        # keeps every line (positions lie in [0, L))
        t_expr = []
        for _ in range(reps):
            t0 = time.perf_counter()
            ing.start_ascii_expr(apath, out[0], n, pos=("$1*1", "$2*1", "$3*1"),
                                 sel="$1 >= 0 && $2 >= 0 && $3 >= 0", comment="#")
            re_ = ing.wait()
            t_expr.append(time.perf_counter() - t0)
        assert re_["n"] == na
    ing.close()
    tf, th = min(t_file), min(t_host)
    ta, te = min(t_ascii), min(t_expr)
    log(f"ingest ASCII: {na} lines, {abytes / 1e6:.0f} MB in {ta * 1e3:.1f} ms "
        f"({abytes / ta / 1e9:.2f} GB/s, {na / ta / 1e6:.1f} M lines/s); with position "
        f"expressions and a selection {te * 1e3:.1f} ms ({na / te / 1e6:.1f} M lines/s)")
    log(f"ingest: file {nbytes / tf / 1e9:.1f} GB/s, host columns {nbytes / th / 1e9:.1f} GB/s; "
        f"{ncat} catalogues ingest+process: sequential {t_seq * 1e3:.1f} ms, "
        f"overlapped {t_ovl * 1e3:.1f} ms")
    return {"workload": f"clustered box catalogue, {n} objects, x/y/z float64 columns ({nbytes / 1e9:.3f} GB)",
            "file_GBps": round(nbytes / tf / 1e9, 2), "host_GBps": round(nbytes / th / 1e9, 2),
            "file_ms": round(tf * 1e3, 2), "stream_catalogues": ncat,
            "sequential_ms": round(t_seq * 1e3, 1), "overlapped_ms": round(t_ovl * 1e3, 1),
            "overlapped_meshes_per_s": round(ncat / t_ovl, 2),
            "ascii": {"lines": na, "bytes": abytes, "ms": round(ta * 1e3, 2),
                      "GBps": round(abytes / ta / 1e9, 3),
                      "lines_per_s": round(na / ta, 1),
                      "expr_sel_ms": round(te * 1e3, 2),
                      "expr_sel_lines_per_s": round(na / te, 1)}}


def parity_vs_oracle(gpu, ref, poles=(0, 2, 4), per_bin_path=None):
    """Per-bin error of the GPU P_l against the FP64 oracle on the same
    catalogue: mode counts compared exactly, |dP_l| / S_b with S_b = |P_0| +
    shot the bin's un-subtracted power, and |dP_0| / |P_0|.  The per-bin
    arrays go to per_bin_path (JSON); the line carries their quantiles."""
    shot = ref["shot"][0]
    has = ref["cnt"] > 0
    S = np.abs(ref["pl"][0][0]) + shot
    out = {"cnt_exact": bool(np.array_equal(gpu["cnt"], ref["cnt"])),
           "nbin": int(len(ref["cnt"]))}
    per_bin = {"k": [float(v) for v in ref["k"]], "cnt": [int(v) for v in ref["cnt"]]}
    for li, ell in enumerate(poles):
        dd = np.where(has, np.abs(gpu["pl"][0][li] - ref["pl"][0][li]) / S, 0.0)
        d = dd[has]
        out[f"max_dP{ell}_over_Sb"] = float(d.max())
        out[f"q50_q90_q99_dP{ell}_over_Sb"] = [float(np.quantile(d, q)) for q in (0.5, 0.9, 0.99)]
        out[f"worst_bin_l{ell}"] = int(np.argmax(dd))
        per_bin[f"dP{ell}_over_Sb"] = [float(f"{v:.3e}") for v in dd]
    if per_bin_path:
        try:
            os.makedirs(os.path.dirname(per_bin_path), exist_ok=True)
            with open(per_bin_path, "w") as f:
                json.dump(per_bin, f)
            out["per_bin_file"] = os.path.relpath(per_bin_path, ROOT)
        except OSError:
            pass
    nz = has & (ref["pl"][0][0] != 0)   # bins whose P_0 is exactly 0 have no relative error
    p0 = ref["pl"][0][0][nz]
    out["max_rel_dP0"] = float(np.max(np.abs(gpu["pl"][0][0][nz] - p0) / np.abs(p0)))
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(box_data, gpu_res=None):
    """FP64 oracle (OpenMP, own FFT) on the whole clustered-box workload: every
    particle, N_g = 1024, PCS interlaced, l = 0,2,4, dk = 0.005 -- one mesh
    (about 10-20 s on the host cores, ~40 GB of host memory)."""
    from oracle import oracle as O
    O.lib()
    conf = O.make_conf(gsize=1024, assign=3, intlace=True, bsize=1000.0, kbin=0.005,
                       poles=[0, 2, 4])
    t = time.perf_counter()
    r = O.run(conf, [box_data])
    wall = time.perf_counter() - t
    ta, tf, tb, tt = r["timings"]
    threads = O.lib().pso_num_threads()
    log(f"cpu oracle, box: {wall:.1f} s wall (assign {ta:.1f}, fft {tf:.1f}, bin {tb:.1f})")
    par = (parity_vs_oracle(gpu_res, r, per_bin_path=os.path.join(
        ROOT, "gpurun_out", "box1024_pcs_parity_per_bin.json")) if gpu_res is not None else None)
    if par:
        log("box vs FP64 oracle: " + ", ".join(f"{k}={v}" for k, v in par.items()))
    return {"value": round(1.0 / wall, 5), "unit": "meshes/s", "cores": threads,
            "cpu_model": cpu_model(),
            "parity_vs_oracle": par,
            "kind": "port",
            "sample": f"oracle/psoracle.c (FP64, OpenMP {threads} threads, own radix-2 FFT) "
                      f"on the full clustered-box workload ({len(box_data)} particles, N_g=1024, PCS "
                      f"interlaced, l=0,2,4): one mesh in {wall:.2f} s (assign {ta:.2f}, "
                      f"FFT {tf:.2f}, binning {tb:.2f})",
            "particles_per_s": round(len(box_data) / wall, 1)}


def _tiles(spec):
    """'tx,ty,tz' -> one tile; 'a,b,c/d,e,f' -> (plain, interlaced) tiles."""
    if not spec:
        return None
    t = [[int(v) for v in s.split(",")] for s in spec.split("/")]
    return t[0] if len(t) == 1 else t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nmock", type=int, default=256)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-survey", action="store_true")
    ap.add_argument("--no-ingest", action="store_true")
    ap.add_argument("--no-alias", action="store_true")
    ap.add_argument("--alias-params", default=None,
                    help="networks of the learned correction (default: the shipped ones)")
    ap.add_argument("--tile", default=None,
                    help="deposit tile of the N_g=1024 box 'tx,ty,tz' instead of autotuning "
                         "(profiling runs); with --only alias 'plain/interlaced' tiles, "
                         "e.g. 8,8,512/8,4,512")
    ap.add_argument("--ysched", default=None,
                    help="y-pass schedule of the N_g=1024 box instead of tuning (profiling runs)")
    ap.add_argument("--batch-tile", default=None,
                    help="deposit tile of the batched workload 'tx,ty,tz' instead of autotuning "
                         "(profiling runs)")
    ap.add_argument("--only", default=None, choices=["batched", "survey", "alias"],
                    help="run one secondary workload alone, no JSON line (profiling runs)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.copy_calib:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); this build "
              "measures the single-GPU path", file=sys.stderr)
        sys.exit(2)
    import powspec_amd as P
    P.load()
    if P.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    arch = P.device_arch()
    if args.only:
        btile = [int(v) for v in args.batch_tile.split(",")] if args.batch_tile else None
        r = {"batched": lambda: run_batched(P, args.nmock, fixed_tile=btile),
             "survey": lambda: run_survey(P),
             "alias": lambda: run_alias(P, params_path=args.alias_params,
                                        fixed_tile=_tiles(args.tile))}[args.only]()
        log(json.dumps(r))
        return
    fixed = [int(v) for v in args.tile.split(",")] if args.tile else None
    data, res, ms_step, stages, n, tile, ysched = run_box(P, args.steps, args.warmup, fixed,
                                                          args.ysched)
    rl = roofline(stages, 1024, 2, n, x_live_fraction(1024, 1000.0, float(res["kedge"][-1])))
    extra = {}
    if not args.no_batched and args.nmock > 0:
        extra = run_batched(P, args.nmock)
    survey = None if args.no_survey else run_survey(P)
    alias = None if args.no_alias else run_alias(P, params_path=args.alias_params)
    ingest = None if args.no_ingest else run_ingest(P, data, tile)
    dropin = None if args.no_ingest else run_dropin(P, data)
    cpu = None if args.no_cpu else cpu_baseline(data, res)
    line = {
        "metric": METRIC,
        "value": round(1e3 / ms_step, 3),
        "unit": "meshes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "clustered box, lognormal Gaussian field 512^3, Poisson "
                               "nbar=0.02 (h/Mpc)^3, Zel'dovich RSD f=0.8; PCS, N_g=1024, "
                               "interlaced, l=0,2,4, dk=0.005, L=1000 Mpc/h",
                   "N_g": 1024, "N_p": n, "assign": "PCS", "interlaced": True,
                   "precision": "FP32 meshes/FFT, FP64 bin sums", "arch": arch,
                   "tile": list(tile), "y_schedule": ysched},
        "particles_per_s": round(n / (ms_step * 1e-3), 1),
        "stage_ms": {k: round(float(v), 4) for k, v in stages.items()},
        "roofline": rl,
        "batched": extra or None,
        "survey": survey,
        "alias_correction": alias,
        "ingest": ingest,
        "dropin_one_shot": dropin,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
