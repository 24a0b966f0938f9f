"""The reference's barrier, contact-matrix and collision-encoding unit suites replayed against
the oracle and the library's host functions (the device replays are in test_gpu_parity.py).
Sources: test/units/simulation_internal/extrusion_barriers_test.cpp:32-93,
test/units/contact_matrix/contact_matrix_dense_test.cpp:68-133,
test/units/simulation_cpu/collision_encoding_test.cpp:25-170."""
import ctypes as C

import numpy as np
import pytest

import unit_kat_replay as K
from modlefork_amd import _lib as L
from modlefork_amd import simulation as S
from modlefork_amd import workloads as W
from oracle import config as OC
from oracle import refcpu as R

KAT = K.load()

BACKENDS = {
    "oracle": (R.lib().rc_stp_active_from_occupancy, R.lib().rc_occupancy_from_stp),
    "library": (L.lib().mgpu_stp_active_from_occupancy, L.lib().mgpu_occupancy_from_stp),
}


@pytest.mark.parametrize("backend", sorted(BACKENDS))
def test_barrier_occupancy_known_answers(backend):
    """occupancy(i) of barriers built from compute_stp_active_from_occupancy gives the occupancy
    back; compute_occupancy_from_stp's six rows, exact where the reference checks equality."""
    to_active, to_occ = BACKENDS[backend]
    for row in KAT["barriers"]["occupancy"]:
        sa = to_active(row["stp_inactive"], row["occupancy"])
        assert K.approx_equal(to_occ(sa, row["stp_inactive"]), row["occupancy"]), row
    for row in KAT["barriers"]["compute_occupancy"]:
        got = to_occ(row["stp_active"], row["stp_inactive"])
        if row["approx"]:
            assert K.approx_equal(got, row["expected"]), (row, got)
        else:
            assert got == row["expected"], (row, got)


def test_barrier_sort_known_answer():
    """The four unsorted barrier positions come out sorted from the library and the oracle."""
    pos = KAT["barriers"]["sort_positions"]
    recs = [W.BedRecord("c", p - 1, p, 0.0, "+") for p in pos]  # midpoint (p - 1 + p + 1) / 2
    want = sorted(pos)
    b = S.prepare_barriers(S.make_params({}), recs, 0, 1000)
    o = OC.prepare_barriers(recs, 0, 1000, OC.transform({}))
    assert b.pos.tolist() == want and list(o.pos) == want


def _oracle_increment(band, missed, nrows, ncols, rows, cols):
    rb = R.RcBand(nrows, ncols, R.ptr(band, C.c_uint32), missed)
    for r, c in zip(rows, cols):
        R.lib().rc_band_increment(C.byref(rb), r, c)
    return int(rb.missed)


@pytest.mark.parametrize("case", ["CMatrix simple", "CMatrix in/decrement"])
def test_contact_matrix_known_answers_oracle(case):
    K.replay_band(KAT["cmatrix"][case], _oracle_increment)


def _oracle_words(idx, ev):
    lib = R.lib()
    w = lib.rc_coll_encode(idx, ev)
    f = 1 if lib.rc_coll_occurred(w, 0) else 0
    for k, kind in enumerate(K.KINDS):
        f |= (1 if lib.rc_coll_occurred(w, kind) else 0) << (1 + k)
        f |= (1 if lib.rc_coll_avoided(w, kind) else 0) << (5 + k)
    return w, lib.rc_coll_index(w), lib.rc_coll_event(w), f


def test_collision_encoding_known_answers_oracle():
    n = K.replay_collisions(KAT["collisions"]["cases"], _oracle_words)
    assert n == 56  # 8 round trips, 8 checks on one word, 40 predicate checks
    # the reference's INDEX_MASK (bits 0..54 of a 64-bit word) round-trips
    assert R.lib().rc_coll_index(R.lib().rc_coll_encode(KAT["collisions"]["index_mask"], 0x11)) \
        == KAT["collisions"]["index_mask"]
