"""Learned alias correction of the plain (non-interlaced) estimator.

A perceptron maps, per k bin of one catalogue, the features
    [k / k_Ny, one-hot assignment scheme (4), 16-band log-power summary of
     the plain spectrum in units of the shot noise over [k_Ny/4, k_Ny],
     log10 of the bin's own plain power over the shot noise]
to the log ratio of the interlaced to the plain (Jing-corrected) monopole,
both un-subtracted (P_0 + shot), so that
    P_corrected(k) = P_plain(k) exp(f(features)).
Network: inputs padded to 32, three hidden layers of 64 with ReLU, one
output (W4 padded to 16 columns); 9.7k live parameters.  Training and
inference run on the GPU with bf16 matrix-core products and FP32
accumulation (psh_mlp_*); `ref_forward` / `ref_grad` are the FP64 CPU
reference of the same network used to pin the GPU gradients.  The
correction applied is the mean of ENSEMBLE networks trained from different
seeds (`train_ensemble`, `predict`).

The correction is a function of the measured plain spectrum, so its floor
is the part of the alias power that the spectrum below k_Ny does not fix:
boxes drawn from one power spectrum with different seeds scatter around
their mean ratio by about 0.5 % at 0.8 k_Ny for P = 1e4 (k/0.1)^-1.5
exp(-(k/0.5)^2) at N_g = 512, and by over 10 % for the highest-variance
lognormal fields of the family (amplitude 2e4, index -1), whose small-scale
power comes from a handful of cells.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

IN, H, OUT = 32, 64, 16
NFEAT = 1 + 4 + 16 + 1
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


def ref_forward(p, X):
    """FP64 forward: (output, cache)."""
    u = {k: v.astype(np.float64) for k, v in unpack(np.asarray(p)).items()}
    x = np.asarray(X, np.float64)
    h1 = np.maximum(x @ u["W1"] + u["b1"], 0)
    h2 = np.maximum(h1 @ u["W2"] + u["b2"], 0)
    h3 = np.maximum(h2 @ u["W3"] + u["b3"], 0)
    y = h3 @ u["W4"] + u["b4"]
    return y[:, 0], (x, h1, h2, h3, u)


def ref_loss(p, X, T):
    y, _ = ref_forward(p, X)
    return float(np.mean((y - T) ** 2))


def ref_grad(p, X, T):
    """FP64 analytic gradient of mean (y - t)^2, in the flat layout."""
    y, (x, h1, h2, h3, u) = ref_forward(p, X)
    B = len(T)
    dy = np.zeros((B, OUT))
    dy[:, 0] = 2 * (y - T) / B
    g = {"W4": h3.T @ dy, "b4": dy.sum(0)}
    d3 = (dy @ u["W4"].T) * (h3 > 0)
    g["W3"], g["b3"] = h2.T @ d3, d3.sum(0)
    d2 = (d3 @ u["W3"].T) * (h2 > 0)
    g["W2"], g["b2"] = h1.T @ d2, d2.sum(0)
    d1 = (d2 @ u["W2"].T) * (h1 > 0)
    g["W1"], g["b1"] = x.T @ d1, d1.sum(0)
    return np.concatenate([g[name].ravel() for name, _ in _SHAPES])


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


# ---- features -------------------------------------------------------------

SUMMARY_LO = 0.25   # the summary covers [SUMMARY_LO k_Ny, k_Ny]


def summary16(k, p_plain, kf, kny, shot):
    """log10 of the plain un-subtracted monopole in units of the shot noise,
    averaged over 16 equal bands of [SUMMARY_LO k_Ny, k_Ny] (bins whose
    centre falls in a band; an empty band takes the nearest populated bin):
    the shape of the spectrum towards the Nyquist frequency, where it sets
    how much power the window folds back, and its level against the shot
    noise.  The bands hold thousands of modes each, so the summary carries
    little sample variance."""
    ok = np.isfinite(p_plain) & (p_plain > 0)
    kk, pp = k[ok], p_plain[ok]
    edges = np.linspace(SUMMARY_LO * kny, kny, 17)
    out = np.empty(16)
    for b in range(16):
        sel = (kk >= edges[b]) & (kk < edges[b + 1])
        if sel.any():
            out[b] = pp[sel].mean()
        else:
            out[b] = pp[np.argmin(np.abs(kk - 0.5 * (edges[b] + edges[b + 1])))]
    return np.log10(out / shot)


def features(k, p_plain_raw, scheme, kf, kny, shot):
    """Rows X[nbin][IN] for one catalogue (bins with finite positive power)."""
    s = summary16(k, p_plain_raw, kf, kny, shot)
    X = np.zeros((len(k), IN), np.float32)
    X[:, 0] = k / kny
    X[:, 1 + scheme] = 1.0
    X[:, 5:21] = s
    with np.errstate(divide="ignore", invalid="ignore"):
        X[:, 21] = np.log10(np.where(p_plain_raw > 0, p_plain_raw, shot) / shot)
    return X


# ---- training data: plain vs interlaced spectra of clustered boxes, N_g = 512 ---

def mock_spectra(P, seed, A, nidx, Ng=512, ng_gen=512, L=1000.0, nbar=0.02, scheme=1,
                 kbin=0.01):
    """(k, raw plain P_0 + shot, raw interlaced P_0 + shot, shot) of one
    clustered box generated on the GPU."""
    dev, n = P.gen_clustered_box(ng=ng_gen, L=L, nbar=nbar, f=0.8, seed=seed, A=A, n=nidx)
    out = []
    for il in (False, True):
        ctx = P.Context(P.make_conf(gsize=Ng, assign=scheme, intlace=il, bsize=L, kbin=kbin,
                                    poles=[0]))
        r = ctx.run_device(dev, n, float(n))
        out.append(r["pl"][0][0] + L ** 3 / n)
        k = r["k"]
        ctx.close()
    dev.free()
    return k, out[0], out[1], L ** 3 / n


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


def dataset(P, seeds, Ng=512, ng_gen=512, L=1000.0, nbar=0.02, scheme=1, kbin=0.01,
            design=None):
    """Feature rows, targets log(P_il / P_plain) and per-mock spectra for the
    mocks `seeds`, mock i with power-law amplitude and index design[i]
    (default: random_design(seeds))."""
    seeds = list(seeds)
    design = random_design(seeds) if design is None else list(design)
    kf, kny = 2 * np.pi / L, np.pi * Ng / L
    Xs, Ts, spectra = [], [], []
    for s, (A, nidx) in zip(seeds, design):
        k, pp, pi, shot = mock_spectra(P, s, A, nidx, Ng, ng_gen, L, nbar, scheme, kbin)
        ok = np.isfinite(pp) & (pp > 0) & (pi > 0)
        X = features(k, pp, scheme, kf, kny, shot)[ok]
        Xs.append(X)
        Ts.append(np.log(pi[ok] / pp[ok]).astype(np.float32))
        spectra.append((k[ok], pp[ok], pi[ok], X))
    return np.concatenate(Xs), np.concatenate(Ts), spectra


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
    loss).  Training is reproducible (fixed-point gradient sums), but which
    minimum a network reaches depends on its seed: on held-out boxes of the
    reference spectrum single networks give worst-bin errors from 0.26 to
    0.71 % by seed, means of four 0.46-0.47 %."""
    ps, ls = zip(*(train(X, T, schedule, batch, seed=s) for s in range(n)))
    return np.stack(ps), float(np.mean(ls))


def predict(params, X):
    """Log correction f(features) per row: one network (params of shape
    [NPAR]) or the mean over an ensemble ([n, NPAR])."""
    P = np.asarray(params, np.float32)
    if P.ndim == 1:
        return gpu_forward(P, X).astype(np.float64)
    return np.mean([gpu_forward(p, X).astype(np.float64) for p in P], axis=0)


def correct(params, k, p_plain_raw, scheme, kf, kny, shot):
    """Corrected un-subtracted monopole P_plain exp(f) for one catalogue
    (bins with non-positive plain power are returned unchanged); params is
    one network or an ensemble."""
    X = features(k, p_plain_raw, scheme, kf, kny, shot)
    y = predict(params, X)
    ok = np.isfinite(p_plain_raw) & (p_plain_raw > 0)
    return np.where(ok, p_plain_raw * np.exp(y), p_plain_raw)


def errors(params, spectra, kny, kmax_frac=0.8):
    """Per catalogue, the worst relative error over k <= kmax_frac k_Ny of
    the corrected and of the Jing-corrected plain monopole against the
    interlaced one: two arrays."""
    e_mlp, e_jing = [], []
    for k, pp, pi, X in spectra:
        sel = k <= kmax_frac * kny
        corr = pp * np.exp(predict(params, X))
        e_mlp.append(float(np.max(np.abs(corr[sel] / pi[sel] - 1))))
        e_jing.append(float(np.max(np.abs(pp[sel] / pi[sel] - 1))))
    return np.array(e_mlp), np.array(e_jing)


def evaluate(params, spectra, kny, kmax_frac=0.8):
    """Worst errors over all catalogues: (corrected, Jing-corrected)."""
    m, j = errors(params, spectra, kny, kmax_frac)
    return float(m.max()), float(j.max())


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
    else:
        print("usage: python -m powspec_amd.mlp bench")
