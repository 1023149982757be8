"""The FFT plans of the library on the CPU (no GPU): the radix plan of every
supported length comes from the library itself (psh_fft_plan), and a model
of the Stockham passes of fft_core.h -- the thread-to-butterfly mapping,
the twiddles, the padded LDS layout and the re/im exchanges between passes
-- run with those radices for the block shapes the z, y and x kernels
launch, must reproduce numpy's FFT, keep every LDS index inside the tile
and, wherever the exchange uses one base plus constant offsets, give the
same address as the element-wise index.  A wrong radix split, stride,
twiddle or padding rule fails here before it reaches a GPU."""
import numpy as np
import pytest

GRIDS_POW2 = [4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048]
GRIDS_MIXED = [48, 96, 192, 384, 768, 1536, 80, 160, 320, 640, 1280]


def _is_pow2(n):
    return n > 0 and n & (n - 1) == 0


def _pow2_floor(n):
    return 1 if n < 2 else 2 * _pow2_floor(n // 2)


def _pow2_ceil(n):
    f = _pow2_floor(n)
    return f if f == n else 2 * f


class _Lds:
    """Padded LDS index of element n of line c: column layout [n][c] with 8
    floats every 8 elements, row layout [c][n] with one float every 16
    elements and an odd line stride."""

    def __init__(self, N, C, colmap):
        self.N, self.C, self.cm = N, C, colmap
        self.LN = N + (N >> 4) + 1
        self.size = N * C + ((N + 7) >> 3) * 8 if colmap else C * self.LN

    def idx(self, c, n):
        return n * self.C + c + ((n >> 3) << 3) if self.cm else c * self.LN + n + (n >> 4)

    def off(self, d):
        return d * self.C + ((d >> 3) << 3) if self.cm else d + (d >> 4)


def _stockham(N, radices, C, T, colmap, inverse=False, seed=0):
    """Run the passes as the kernels do; returns (max relative error against
    numpy, separable-address violations, largest values per thread)."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((C, N)) + 1j * rng.standard_normal((C, N))
    L = _Lds(N, C, colmap)
    tw = np.exp(-2j * np.pi * np.arange(N) / N)
    ns = [1]
    for r in radices[:-1]:
        ns.append(ns[-1] * r)

    def geom(p):
        R = radices[p]
        NBF = N // R
        NBT = C * NBF
        return R, ns[p], NBF, NBT, (NBT + T - 1) // T

    def mapping(p, t, q):
        R, NS, NBF, NBT, NB = geom(p)
        b = t + q * T
        c, j = (b % C, b // C) if colmap else (b // NBF, b % NBF)
        return b < NBT, c, j

    R, NS, NBF, NBT, NB = geom(0)
    vals = {}
    for t in range(T):
        for q in range(NB):
            ok, c, j = mapping(0, t, q)
            vals[t, q] = [X[c, j + r * NBF] if ok else 0j for r in range(R)]
    bad, vmax = 0, 0
    for p in range(len(radices)):
        R, NS, NBF, NBT, NB = geom(p)
        vmax = max(vmax, NB * R)
        Wr = np.exp((1 if inverse else -1) * 2j * np.pi * np.outer(np.arange(R), np.arange(R)) / R)
        for t in range(T):
            for q in range(NB):
                ok, c, j = mapping(p, t, q)
                v = np.array(vals[t, q])
                if p > 0:
                    k, step = j % NS, N // (NS * R)
                    w = tw[np.arange(R) * k * step]
                    v = v * (np.conj(w) if inverse else w)
                vals[t, q] = list(Wr @ v)
        if p + 1 == len(radices):
            break
        R1, NS1, NBF1, NBT1, NB1 = geom(p + 1)
        blk = 8 if colmap else 16
        sep0 = _is_pow2(NS) or NS % blk == 0
        sep1 = _is_pow2(NBF1) or NBF1 % blk == 0
        parts = []
        for im in (0, 1):
            lds = np.full(L.size, np.nan)
            for t in range(T):
                for q in range(NB):
                    ok, c, j = mapping(p, t, q)
                    if not ok:
                        continue
                    e0 = (j // NS) * NS * R + j % NS
                    for r in range(R):
                        a = L.idx(c, e0 + r * NS)
                        bad += sep0 and a != L.idx(c, e0) + L.off(r * NS)
                        assert 0 <= a < L.size
                        lds[a] = vals[t, q][r].imag if im else vals[t, q][r].real
            got = {}
            for t in range(T):
                for q in range(NB1):
                    ok, c, j = mapping(p + 1, t, q)
                    row = []
                    for r in range(R1):
                        if not ok:
                            row.append(0.0)
                            continue
                        a = L.idx(c, j + r * NBF1)
                        bad += sep1 and a != L.idx(c, j) + L.off(r * NBF1)
                        assert 0 <= a < L.size and not np.isnan(lds[a])
                        row.append(lds[a])
                    got[t, q] = row
            parts.append(got)
        vals = {k: [complex(a, b) for a, b in zip(parts[0][k], parts[1][k])] for k in parts[0]}
    R, NS, NBF, NBT, NB = geom(len(radices) - 1)
    assert NS * R == N
    Y = np.zeros((C, N), complex)
    for t in range(T):
        for q in range(NB):
            ok, c, j = mapping(len(radices) - 1, t, q)
            if ok:
                e0 = (j // NS) * NS * R + j % NS
                for r in range(R):
                    Y[c, e0 + r * NS] = vals[t, q][r]
    ref = np.fft.ifft(X, axis=1) * N if inverse else np.fft.fft(X, axis=1)
    return float(np.abs(Y - ref).max() / np.abs(ref).max()), bad, vmax


def test_plans_factor_every_grid(psh_host):
    psh = psh_host
    for N in GRIDS_POW2 + GRIDS_MIXED:
        odd = N // (N & -N)
        for plan in (0, 1, 2):
            r = psh.fft_plan(N, plan)
            assert int(np.prod(r)) == N, (N, plan, r)
            assert all(x in (2, 4, 8, 16) for x in r[: len(r) - (odd > 1)]), (N, plan, r)
            if odd > 1:
                assert r[-1] == odd
        # plan 0: the fewest passes of radix <= 16, larger radices first
        r0 = [x for x in psh.fft_plan(N, 0) if x in (2, 4, 8, 16)]
        assert len(r0) == -(-int(np.log2(N & -N)) // 4) and r0 == sorted(r0, reverse=True)
    assert psh.fft_plan(100) == [] and psh.fft_plan(112) == [] and psh.fft_plan(3) == []


# the block shapes of the kernels: y pass (16 columns, 32 or 16 values per
# thread; 8 and 32 columns), x pass (16 or 8 columns, 16 or 8 values), z pass
# (rows of N/2 points, 4096 / (N/2) rows rounded down to a power of two, 256
# threads); thread counts are powers of two, rounded up from C N / values
SHAPES = [(16, 32), (16, 16), (8, 16), (32, 32), (16, 8)]


@pytest.mark.parametrize("N", [16, 64, 256, 48, 80, 96, 160, 192, 320, 384])
def test_stockham_model_matches_numpy(psh_host, N):
    psh = psh_host
    for plan in (0, 1, 2):
        radices = psh.fft_plan(N, plan)
        for C, vpt in SHAPES:
            T = max(64, min(1024, _pow2_ceil(C * N // vpt)))
            err, bad, _ = _stockham(N, radices, C, T, True, seed=N + plan)
            assert err < 1e-12 and bad == 0, (N, plan, C, T, err, bad)
    M = N // 2
    C = _pow2_floor(4096 // M)
    for inverse in (False, True):
        err, bad, _ = _stockham(M, psh.fft_plan(M, 0), C, 256, False, inverse=inverse)
        assert err < 1e-12 and bad == 0, (M, err, bad)
