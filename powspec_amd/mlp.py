"""Learned alias correction of the plain (non-interlaced) box estimator.

The library stage (psh_ctx_set_mu_cells, psh_ctx_set_alias_mlp) bins a plain
box run in (k bin, |mu| cell) pairs and corrects every cell's raw sums by
exp(f), f the mean output of a few perceptrons evaluated on the device on
feature rows the library builds from the run itself (psh_ctx_alias_rows):
    [k / k_Ny, |mu| cell centre, one-hot scheme (4), 16-band log10 summary
     of the un-subtracted monopole over the shot noise on [k_Ny/4, k_Ny],
     the bin's own log10(P_0 / shot), log10 of the cell's share of the bin]
so that P_l of every pole is corrected, not only the monopole.  The target
of a cell is log(S_il / S_plain), the ratio of its monopole sums in an
interlaced and a plain run of the same catalogue.  Network: inputs padded
to 32, three hidden layers of 64 with ReLU, one output (W4 padded to 16
columns); 10.0k live parameters, bf16 matrix-core products with FP32
accumulation for training (psh_mlp_train) and inference.

This module drives the training on the GPU: catalogues from the library's
own generator, plain and interlaced runs, feature rows and targets from the
library, and evaluation against the interlaced estimator, which stays the
fallback and the judge of the error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

IN, H, OUT = 32, 64, 16
NFEAT = 24
NMU = 8
_SHAPES = [("W1", (IN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)),
           ("b3", (H,)), ("W4", (H, OUT)), ("b4", (OUT,))]
NPAR = sum(int(np.prod(s)) for _, s in _SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shape in _SHAPES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def init_params(seed=0):
    """He-scaled weights on the live entries; padding (W1 rows >= NFEAT, W4
    columns >= 1, b4[1:]) zero and kept zero by training."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPAR, np.float32)
    u = unpack(p)
    u["W1"][:NFEAT] = rng.normal(0, np.sqrt(2.0 / NFEAT), (NFEAT, H))
    u["W2"][:] = rng.normal(0, np.sqrt(2.0 / H), (H, H))
    u["W3"][:] = rng.normal(0, np.sqrt(2.0 / H), (H, H))
    u["W4"][:, 0] = rng.normal(0, np.sqrt(1.0 / H), H)
    return p


# ---- GPU ------------------------------------------------------------------

def _lib():
    from powspec_amd import load
    return load()


def _dev(a):
    from powspec_amd import DeviceArray
    return DeviceArray.from_numpy(np.ascontiguousarray(a, np.float32))


def gpu_forward(p, X):
    from powspec_amd import PshError
    d_p, d_x = _dev(p), _dev(X)
    d_y = _dev(np.zeros(len(X), np.float32))
    rc = _lib().psh_mlp_forward(d_p.ptr, d_x.ptr, len(X), d_y.ptr)
    if rc:
        raise PshError(rc, "psh_mlp_forward failed")
    return d_y.to_numpy((len(X),), np.float32)


def gpu_grad(p, X, T):
    from powspec_amd import PshError
    d_p, d_x, d_t = _dev(p), _dev(X), _dev(T)
    d_g = _dev(np.zeros(NPAR, np.float32))
    loss = C.c_float()
    rc = _lib().psh_mlp_grad(d_p.ptr, d_x.ptr, d_t.ptr, len(T), d_g.ptr, C.byref(loss))
    if rc:
        raise PshError(rc, "psh_mlp_grad failed")
    return d_g.to_numpy((NPAR,), np.float32), loss.value


def gpu_train(p, X, T, epochs=200, batch=256, lr=2e-3, seed=0):
    """Adam on the GPU; the host shuffles once per call.  Returns (params,
    final loss)."""
    from powspec_amd import PshError
    order = np.random.default_rng(seed).permutation(len(T))
    d_p, d_x, d_t = _dev(p), _dev(np.asarray(X)[order]), _dev(np.asarray(T)[order])
    loss = C.c_float()
    rc = _lib().psh_mlp_train(d_p.ptr, d_x.ptr, d_t.ptr, len(T), int(batch), int(epochs),
                              C.c_float(lr), C.c_uint(seed), C.byref(loss))
    if rc:
        raise PshError(rc, "psh_mlp_train failed")
    return d_p.to_numpy((NPAR,), np.float32), loss.value


# ---- training data: plain vs interlaced runs of clustered boxes ------------

def box_conf(P, Ng=512, L=1000.0, kbin=0.01, scheme=1, intlace=False):
    return P.make_conf(gsize=Ng, assign=scheme, intlace=intlace, bsize=L, kbin=kbin,
                       poles=[0, 2, 4])


def mock_runs(P, seed, A, nidx, Ng=512, ng_gen=512, L=1000.0, nbar=0.02, scheme=1,
              kbin=0.01, nmu=NMU):
    """One clustered box generated on the GPU, run plain and interlaced with
    |mu| cells: (feature rows, plain cells, interlaced cells, interlaced
    P_l, plain P_l, shot), cells [3][nbin][nmu]."""
    dev, n = P.gen_clustered_box(ng=ng_gen, L=L, nbar=nbar, f=0.8, seed=seed, A=A, n=nidx)
    out = {}
    for il in (False, True):
        ctx = P.Context(box_conf(P, Ng, L, kbin, scheme, il))
        ctx.set_mu_cells(nmu)
        r = ctx.run_device(dev, n, float(n))
        out[il] = (ctx.cells(), r["pl"][0])
        if not il:
            out["X"] = ctx.alias_rows()
            out["k"] = r["k"]
        ctx.close()
    dev.free()
    return out["X"], out[False][0], out[True][0], out[True][1], out[False][1], L ** 3 / n


def targets(cells_plain, cells_il, X):
    """log(S_il / S_plain) of the monopole cell sums; rows without features
    or power are left out (mask)."""
    sp, si = cells_plain[0].ravel(), cells_il[0].ravel()
    ok = (X[:, 0] != 0) & (sp > 0) & (si > 0)
    t = np.zeros(len(sp), np.float32)
    t[ok] = np.log(si[ok] / sp[ok])
    return t, ok


def grid_design(m=8, amp=(5e3, 2e4), index=(-2.0, -1.0)):
    """m x m training design over the power-law family: amplitudes
    log-spaced and indices evenly spaced over the closed ranges, so that any
    member of the family lies inside the design."""
    return [(a, n) for a in np.geomspace(*amp, m) for n in np.linspace(*index, m)]


def random_design(seeds, rng_seed=0, amp=(5e3, 2e4), index=(-2.0, -1.0)):
    """(A, n) drawn uniformly from the ranges, one pair per seed."""
    out = []
    for s in seeds:
        r = np.random.default_rng(rng_seed + s)
        out.append((r.uniform(*amp), r.uniform(*index)))
    return out


def dataset(P, seeds, design=None, **kw):
    """Training rows, targets and per-mock evaluation records for the
    mocks `seeds` (mock i with power-law amplitude and index design[i],
    default random_design(seeds))."""
    seeds = list(seeds)
    design = random_design(seeds) if design is None else list(design)
    Xs, Ts, recs = [], [], []
    for s, (A, nidx) in zip(seeds, design):
        X, cp, ci, pl_il, pl_plain, shot = mock_runs(P, s, A, nidx, **kw)
        t, ok = targets(cp, ci, X)
        Xs.append(X[ok])
        Ts.append(t[ok])
        recs.append({"seed": s, "A": A, "n": nidx, "pl_il": pl_il, "pl_plain": pl_plain,
                     "shot": shot})
    return np.concatenate(Xs), np.concatenate(Ts), recs


SCHEDULE = ((300, 2e-3), (300, 5e-4), (200, 1e-4))


def train(X, T, schedule=SCHEDULE, batch=256, seed=0):
    """Adam on the GPU from He-initialised weights, in stages of (epochs,
    learning rate) with the moments restarted at each stage."""
    p = init_params(seed)
    loss = float("nan")
    for i, (ep, lr) in enumerate(schedule):
        p, loss = gpu_train(p, X, T, epochs=ep, batch=batch, lr=lr, seed=seed + 17 * i)
    return p, loss


ENSEMBLE = 4


def train_ensemble(X, T, n=ENSEMBLE, schedule=SCHEDULE, batch=256):
    """n networks trained from seeds 0..n-1: (params [n, NPAR], mean final
    loss).  Training is reproducible (fixed-point gradient sums); which
    minimum a network reaches depends on its seed, and the mean of a few
    averages that spread out."""
    ps, ls = zip(*(train(X, T, schedule, batch, seed=s) for s in range(n)))
    return np.stack(ps), float(np.mean(ls))


# ---- evaluation of the library stage against the interlaced estimator ------

def corrected_runs(P, params, seeds, design, Ng=512, ng_gen=512, L=1000.0, nbar=0.02,
                   scheme=1, kbin=0.01, nmu=NMU):
    """Per held-out catalogue: (k, corrected P_l of the library stage,
    interlaced P_l, plain Jing-corrected P_l, shot)."""
    out = []
    for s, (A, nidx) in zip(seeds, design):
        dev, n = P.gen_clustered_box(ng=ng_gen, L=L, nbar=nbar, f=0.8, seed=s, A=A, n=nidx)
        plain = P.Context(box_conf(P, Ng, L, kbin, scheme, False))
        plain.set_mu_cells(nmu)
        jing = plain.run_device(dev, n, float(n))
        plain.set_alias_mlp(params)
        corr = plain.run_device(dev, n, float(n))
        il = P.Context(box_conf(P, Ng, L, kbin, scheme, True)).run_device(dev, n, float(n))
        dev.free()
        out.append((corr["k"], corr["pl"][0], il["pl"][0], jing["pl"][0], L ** 3 / n))
    return out


def errors(runs, kny, kmax_frac=0.8):
    """Worst error per catalogue over k <= kmax_frac k_Ny, for each pole,
    relative to the bin's un-subtracted interlaced monopole S_b = P_0 + shot
    (the parity scale): arrays [ncat][3] for the library stage and for the
    Jing-corrected plain estimator."""
    e_c, e_j = [], []
    for k, pc, pi, pj, shot in runs:
        sel = k <= kmax_frac * kny
        S = pi[0][sel] + shot
        e_c.append([float(np.max(np.abs(pc[l][sel] - pi[l][sel]) / S)) for l in range(3)])
        e_j.append([float(np.max(np.abs(pj[l][sel] - pi[l][sel]) / S)) for l in range(3)])
    return np.array(e_c), np.array(e_j)


# ---- shipped networks ---------------------------------------------------------

PARAMS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "alias_cic_ng512.npz")


def save_params(params, path=PARAMS_FILE, **meta):
    np.savez(path, params=np.asarray(params, np.float32),
             **{k: np.asarray(v) for k, v in meta.items()})


def load_params(path=PARAMS_FILE):
    """The shipped ensemble [n, NPAR] (trained by `python -m powspec_amd.mlp
    train`), or None when absent."""
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as z:
        return np.asarray(z["params"], np.float32)


def train_and_save(P, path=PARAMS_FILE):
    """The shipped networks: 64 clustered boxes (CIC, N_g = 512, seeds
    5000..5063) on the 8 x 8 design of the power-law family, four networks."""
    X, T, _ = dataset(P, range(5000, 5064), design=grid_design(8))
    print(f"training rows: {len(T)}", flush=True)
    ps, ls = [], []
    for s in range(ENSEMBLE):
        p, l = train(X, T, seed=s)
        print(f"network {s}: final loss {l:.3e}", flush=True)
        ps.append(p)
        ls.append(l)
    params, loss = np.stack(ps), float(np.mean(ls))
    save_params(params, path, rows=len(T), loss=loss, nmu=NMU, scheme=1, ng=512,
                seeds=np.arange(5000, 5064))
    return params, loss


def _bench():
    """Kernel workload for profiling: inference on 2^20 feature rows (5 calls),
    training on 10240 rows (batch 256, 50 epochs), one gradient over 65536
    rows; prints the wall time of each part."""
    import time
    rng = np.random.default_rng(0)
    B = 1 << 20
    X = np.zeros((B, IN), np.float32)
    X[:, :NFEAT] = rng.normal(0, 1, (B, NFEAT))
    p = init_params(0)
    t = time.perf_counter()
    for _ in range(5):
        gpu_forward(p, X)
    print(f"forward 5 x {B} rows: {time.perf_counter() - t:.3f} s (incl. transfers)")
    Xt = X[:10240]
    T = (0.1 * np.sin(Xt[:, 0])).astype(np.float32)
    t = time.perf_counter()
    _, loss = gpu_train(p, Xt, T, epochs=50, batch=256, lr=1e-3)
    print(f"train 10240 rows x 50 epochs: {time.perf_counter() - t:.3f} s, loss {loss:.3e}")
    gpu_grad(p, X[:65536], np.zeros(65536, np.float32))


if __name__ == "__main__":
    import sys
    if sys.argv[1:] == ["bench"]:
        _bench()
    elif sys.argv[1:2] == ["train"]:
        import powspec_amd as _P
        _P.load()
        out = sys.argv[2] if len(sys.argv) > 2 else PARAMS_FILE
        _, loss = train_and_save(_P, out)
        print(f"trained networks written to {out} (mean final loss {loss:.3e})")
    else:
        print("usage: python -m powspec_amd.mlp bench | train [path]")
