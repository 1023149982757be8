"""The libast golden vectors of the ASCII ingest's expressions
(tests/golden/ascii_expr_vectors.json, from the reference's lib/libast.c)
against what the grammar fixes in closed form: left-associative **, unary
minus binding tighter than **, fmod for %, double division of integer
columns in value trees, integer comparisons in selections."""
import json
import math
import os

import numpy as np

V = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                "ascii_expr_vectors.json")))
C = np.array([[float(x) for x in r] for r in V["rows"]])
BY = {(c["kind"], c["expr"]): c for c in V["cases"]}


def _vals(e):
    return np.array(BY["D", e]["values"], dtype=float)


def test_fixture_shape():
    assert V["formatter"] == "%lf %d %lf %lf %d %lf" and C.shape == (160, 6)
    assert all(r[1].lstrip("-").isdigit() and r[4].lstrip("-").isdigit() for r in V["rows"])


def test_grammar_closed_forms():
    assert np.all(_vals("2**3**2") == 64.0)                 # (2**3)**2, not 2**9
    assert np.array_equal(_vals("-$1**2"), C[:, 0] ** 2)    # (-$1)**2
    assert np.array_equal(_vals("$1"), C[:, 0]) and np.array_equal(_vals("${3}"), C[:, 2])
    assert np.array_equal(_vals("$1-$3-$4"), (C[:, 0] - C[:, 2]) - C[:, 3])
    assert np.array_equal(_vals("$1/$3/$4"), (C[:, 0] / C[:, 2]) / C[:, 3])
    assert np.array_equal(_vals("$1 % 0.7"), [math.fmod(x, 0.7) for x in C[:, 0]])
    assert np.array_equal(_vals("-$4 % 0.3"), [math.fmod(-x, 0.3) for x in C[:, 3]])
    with np.errstate(divide="ignore", invalid="ignore"):   # double division: 0/0, n/0
        assert np.array_equal(_vals("$2/$5"), C[:, 1] / C[:, 4], equal_nan=True)
    assert np.array_equal(_vals("- -$3"), C[:, 2])
    assert np.array_equal(_vals("abs(-$1)*-2"), np.abs(C[:, 0]) * -2)
    assert np.array_equal(np.isnan(_vals("sqrt($3)")), C[:, 2] < 0)


def test_selection_closed_forms():
    def keep(e):
        return np.array(BY["B", e]["keep"])
    assert np.array_equal(keep("$2 == 4"), C[:, 1] == 4)
    assert np.array_equal(keep("!$2"), C[:, 1] == 0)
    assert np.array_equal(keep("-$2 < -3"), C[:, 1] > 3)
    assert np.array_equal(keep("$1 > 0.2 && $3 < 0.5"), (C[:, 0] > 0.2) & (C[:, 2] < 0.5))
    # && binds tighter than ||
    assert np.array_equal(keep("$4 > 0 && $4 < 0.5 || $1 > 0.9"),
                          ((C[:, 3] > 0) & (C[:, 3] < 0.5)) | (C[:, 0] > 0.9))
    # comparisons bind tighter than ==
    assert np.array_equal(keep("$1 > 0.3 == $3 > 0.3"), (C[:, 0] > 0.3) == (C[:, 2] > 0.3))
    assert np.array_equal(keep("isfinite(ln($6))"), C[:, 5] > 0)
    # long arithmetic: C's truncating division and remainder, integer power
    assert np.array_equal(keep("$2 % 2 == 0"), C[:, 1] % 2 == 0)
    assert np.array_equal(keep("$5 / 2 == -1"), np.trunc(C[:, 4] / 2) == -1)
    assert np.array_equal(keep("-$5 % 3 == -1"), np.fmod(-C[:, 4], 3) == -1)
    assert np.array_equal(keep("2 ** $2 >= 64"), C[:, 1] >= 6)
    assert np.array_equal(keep("$2 ** 63 != 0"), C[:, 1] == 1)   # overflow: only 1 ** n
    assert np.array_equal(keep("(-1) ** ($2 - 9) == -1"), C[:, 1] % 2 == 0)
    # bitwise operators of longs; == binds tighter than &
    i2, i5 = C[:, 1].astype(np.int64), C[:, 4].astype(np.int64)
    assert np.array_equal(keep("($2 & 1) == 1"), (i2 & 1) == 1)
    assert np.array_equal(keep("~$2 < -5"), ~i2 < -5)
    assert np.array_equal(keep("$5 >> 1 == -1"), (i5 >> 1) == -1)
    wrapped = (i2.astype(np.uint64) << np.uint64(62)).astype(np.int64)   # two's complement
    assert np.array_equal(keep("$2 << 62 < 0"), wrapped < 0)
    assert not BY["B", "$2 & 1 == 1"]["builds"]     # $2 & (1 == 1): a long and a bool
    assert BY["B", "($1 & 1) == 1"]["eval_error"]   # a double column


def test_refusals_recorded():
    assert not BY["D", "$1 < 2"]["builds"] and not BY["D", "$12"]["builds"]
    assert not BY["B", "$1"]["builds"] and BY["B", "$1 > 0.5 && $2"]["eval_error"]
    assert V["refused_here"] == []
