"""Writes ascii_expr_vectors.json: rows of a six-column ASCII catalogue
(formatter "%lf %d %lf %lf %d %lf") and, for each expression, the value the
reference's own libast gives per row (oracle/_ref/ast_ref: lib/libast.c
compiled where it lies, driven as io/read_ascii.c:233-340 drives it).

Run from the repository root after `make -C oracle ref`:
    python tests/golden/make_ascii_expr_vectors.py
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
TYPES = "d i d d i d"
FMTR = "%lf %d %lf %lf %d %lf"

# (kind, expression): D value trees, B selections
EXPRS = [
    ("D", "$1"), ("D", "${3}"), ("D", "$1*2+${3}"), ("D", "-$1**2"), ("D", "2**3**2"),
    ("D", "$1-$3-$4"), ("D", "$1/$3/$4"), ("D", "($1+$3)*$4"), ("D", "$1 % 0.7"),
    ("D", "-$4 % 0.3"), ("D", "abs($3-$4)"), ("D", "sqrt(abs($1))"), ("D", "ln(abs($4)+1)"),
    ("D", "log(abs($3)+1e-3)"), ("D", "$2*0.5+$5"), ("D", "$2/$5"), ("D", "$2 % 3"),
    ("D", "$6**0.5"), ("D", "1e3*$1"), ("D", ".5*$1"), ("D", "-(-$3)"), ("D", "- -$3"),
    ("D", "abs(-$1)*-2"), ("D", "$1**2+$3**2+$4**2"), ("D", "sqrt($1**2 + $3**2)"),
    ("D", "sqrt($3)"), ("D", "ln($4)"), ("D", "2*$1-$3/4+$4*$6"),
    ("D", "abs($4)+0.5"), ("D", "1/(1+$6*$2)"), ("D", "$6*1e-4"),
    ("D", "((($1))))"), ("D", "$1 +"), ("D", "$1 < 2"), ("D", "$0"), ("D", "$12"),
    ("D", "${7}"), ("D", "inf"), ("D", "sqrt $1"), ("D", "$1 $3"), ("D", "1e400*$1"),
    ("D", "!$1"), ("D", "$2 & 1"),
    ("B", "$1 > 0.2"), ("B", "$1 > 0.2 && $3 < 0.5"), ("B", "$1 < 0 || $4 >= 0.5"),
    ("B", "!($1 > 0.5)"), ("B", "$2 >= 2 && $5 != 3"), ("B", "$2 == 4"),
    ("B", "isfinite(ln($6))"), ("B", "!$2"), ("B", "$1*$3 > 0.1"), ("B", "abs($3) <= 0.25"),
    ("B", "$2 + $5 > 6"), ("B", "$2 * 2 < 7"), ("B", "-$2 < -3"),
    ("B", "sqrt($6) > 0.5 || $2 == 0"), ("B", "$1 > 0.3 == $3 > 0.3"),
    ("B", "$4 > 0 && $4 < 0.5 || $1 > 0.9"), ("B", "!!($3 < 0)"), ("B", "$1 > 0"),
    ("B", "$1"), ("B", "$1 + 1"), ("B", "($1 > 0.5) + 1 > 0"), ("B", "$1 > 0.5 && $2"),
    ("B", "$1 > (0.5"),
    # libast's long arithmetic: truncating division, C remainder, integer power
    ("B", "$2 / 2 > 1"), ("B", "$2 % 2 == 0"), ("B", "-$5 % 3 == -1"), ("B", "$5 / 2 == -1"),
    ("B", "$2 ** 2 > 10"), ("B", "$5 ** 3 < -8"), ("B", "2 ** $2 >= 64"),
    ("B", "$5 ** 300 == 0"), ("B", "(-1) ** ($2 - 9) == -1"), ("B", "$2 ** 63 != 0"),
    # bitwise operators of longs (x86-64 shifts: counts modulo 64)
    ("B", "($2 & 1) == 1"), ("B", "$2 << 1 > 2"), ("B", "~$2 < -5"), ("B", "($2 | $5) > 3"),
    ("B", "($2 ^ 5) == 3"), ("B", "$5 >> 1 == -1"), ("B", "1 << $2 >= 32"),
    ("B", "$2 << 62 < 0"), ("B", "$2 & 1 == 1"), ("B", "($1 & 1) == 1"),
]
# accepted by libast, refused here
REFUSED_HERE = []


def rows(n=160, seed=20261203):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        a = rng.uniform(-0.2, 1.0)
        b = int(rng.integers(0, 9))
        c = rng.uniform(-0.5, 0.8)
        d = rng.normal(0.3, 0.4)
        e = int(rng.integers(-4, 6))
        f = rng.uniform(0.0, 2.0) if i % 7 else 0.0
        out.append([repr(float(a)), str(b), repr(float(c)), repr(float(d)), str(e),
                    repr(float(f))])
    return out


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "ast_ref")
    if not os.path.exists(exe):
        sys.exit("build oracle/_ref/ast_ref first: make -C oracle ref")
    R = rows()
    inp = ["6", TYPES, str(len(EXPRS))] + [f"{k} {e}" for k, e in EXPRS]
    inp += [str(len(R))] + [" ".join(r) for r in R]
    res = subprocess.run([exe], input="\n".join(inp) + "\n", capture_output=True, text=True,
                         check=True).stdout.split("\n")
    built = res[:len(EXPRS)]
    vals = [line.split(" ") for line in res[len(EXPRS):len(EXPRS) + len(R)]]
    cases = []
    for j, (k, e) in enumerate(EXPRS):
        c = {"kind": k, "expr": e, "builds": built[j] == "OK"}
        if c["builds"]:
            col = [v[j] for v in vals]
            if k == "D":
                c["values"] = [float.fromhex(x) if x not in ("E", "-") else None for x in col]
            else:
                c["keep"] = [x == "1" for x in col]
                c["eval_error"] = any(x in ("E", "-") for x in col)
        cases.append(c)
    out = {"formatter": FMTR, "rows": R, "cases": cases, "refused_here": REFUSED_HERE,
           "source": "lib/libast.c of the reference, via oracle/ast_ref.c"}
    with open(os.path.join(HERE, "ascii_expr_vectors.json"), "w") as fh:
        json.dump(out, fh, separators=(",", ":"))
    print(f"{len(R)} rows, {len(cases)} expressions "
          f"({sum(c['builds'] for c in cases)} build in libast)")


if __name__ == "__main__":
    main()
