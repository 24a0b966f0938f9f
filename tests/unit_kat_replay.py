"""Replays the known answers of the reference's barrier, contact-matrix and collision-encoding
unit suites (tests/golden/reference_unit_kat.json, written by
tests/golden/extract_reference_unit_kat.py) through a backend: the oracle, the library's host
functions or the device.

Contact matrix: `increment(band, missed, nrows, ncols, rows, cols) -> missed` is the backend's;
the suites' decrements (the reference matrix also subtracts, the simulation never does) and
their `get` checks read the band through the ContactMatrixDense layout, pixel (i, j) with
i <= j at data[j * nrows + (j - i)] (contact_matrix_internal_impl.hpp:19-42).
Collisions: `words(idx, ev) -> (word, index, event, flags)` with flags bit 0 the COLLISION bit,
bits 1-4 collision_occurred(CHROM_BOUNDARY, LEF_BAR, LEF_LEF_PRIMARY, LEF_LEF_SECONDARY), bits
5-8 collision_avoided of the same events; collision_avoided() without an event is the reference's
`!collision_occurred() && word != 0` (collision_encoding_impl.hpp:226-228)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "reference_unit_kat.json")
KINDS = [0x08, 0x04, 0x02, 0x01]  # CHROM_BOUNDARY, LEF_BAR, LEF_LEF_PRIMARY, LEF_LEF_SECONDARY


def load():
    with open(GOLDEN) as fh:
        return json.load(fh)


def approx_equal(a, b):
    """Catch::Approx's default test: |a - b| <= 100 float epsilons of max(|a|, |b|)."""
    return abs(a - b) <= 100 * np.finfo(np.float32).eps * max(abs(a), abs(b))


def pixel_index(nrows, i, j):
    i, j = min(i, j), max(i, j)
    return j * nrows + (j - i)


def replay_band(case, increment):
    nrows, ncols = case["nrows"], case["ncols"]
    band = np.zeros(nrows * ncols + 1, np.uint32)
    missed = 0
    for op in case["ops"]:
        kind = op[0]
        if kind == "increment":
            for _ in range(op[3]):
                missed = increment(band, missed, nrows, ncols, [op[1]], [op[2]])
        elif kind == "decrement":
            band[pixel_index(nrows, op[1], op[2])] -= op[3]
        elif kind == "get":
            assert band[pixel_index(nrows, op[1], op[2])] == op[3], op
        elif kind == "tot":
            assert int(band.sum()) == op[1], op
        elif kind == "missed":
            assert missed == op[1], op
        else:
            raise ValueError(op)
    return band, missed


def _check(word, flags, what, kind, expected):
    if kind == 0:
        got = bool(flags & 1) if what == "occurred" else (not flags & 1) and word != 0
    else:
        bit = (1 if what == "occurred" else 5) + KINDS.index(kind)
        got = bool((flags >> bit) & 1)
    return got == expected


def replay_collisions(cases, words):
    """Runs every case; returns the number of checks made."""
    nchecks = 0
    for name, ops in cases.items():
        idx, ev = 0, 0
        for op in ops:
            kind = op[0]
            if kind == "roundtrip":
                w, i, e, _ = words(op[1], op[2])
                assert (i, e) == (op[1], op[2]), (name, op, hex(w))
            elif kind == "check":
                w, _, _, f = words(op[1], op[2])
                assert _check(w, f, op[3], op[4], op[5]), (name, op, hex(w), bin(f))
            elif kind == "set_event":
                ev = op[1]
                continue
            elif kind == "set_idx":
                idx |= op[1]
                continue
            elif kind == "set":
                idx, ev = op[1], op[2]
                continue
            elif kind == "word_check":
                w, _, _, f = words(idx, ev)
                assert _check(w, f, op[1], op[2], op[3]), (name, op, hex(w), bin(f))
            elif kind == "word_index":
                _, i, _, _ = words(idx, ev)
                assert i == op[1], (name, op)
            else:
                raise ValueError(op)
            nchecks += 1
    return nchecks


def _value(expr: str, i: int, j: int) -> int:
    """An extracted fill expression: integers, i, j, +, - and parentheses only."""
    if not set(expr) <= set("0123456789ij+-() "):
        raise ValueError(f"unexpected fill expression {expr!r}")
    return int(eval(expr, {"__builtins__": {}}, {"i": i, "j": j}))


def block_band(case):
    """The band of "CMatrix get w/ block": each fill sets pixel (i, j) over its loop ranges (j
    from i when the loop says so) to its expression."""
    nrows, ncols = case["nrows"], case["ncols"]
    band = np.zeros(nrows * ncols + 1, np.uint32)
    for f in case["fills"]:
        for i in range(f["i"][0], f["i"][1]):
            j0 = i if f["j"][0] == "i" else f["j"][0]
            for j in range(j0, f["j"][1]):
                band[pixel_index(nrows, i, j)] = _value(f["value"], i, j)
    return band
