"""The oracle's optional per-cell outputs: the 1D LEF occupancy track
(register_1d_lef_occupancy, register_contacts.cpp:200-232) and the model internal-state log
(Simulation::dump_stats, simulation.cpp:1166-1224)."""
import json
import os

import numpy as np

from oracle import config as OC
from oracle import refcpu as R
from tests import kat_replay as KR
from tests.test_oracle_sim import setup, tiny

HERE = os.path.dirname(os.path.abspath(__file__))


def _occ_len(ch, bin_size):
    return (ch.size + bin_size - 1) // bin_size


def test_occupancy_adds_both_units_of_every_sampling_event():
    rc, ch, bars, nl, t, te, nrows, ncols = setup(tiny())
    occ = np.zeros(_occ_len(ch, rc["bin_size"]), np.uint64)
    band, missed, st = R.simulate_cells(rc, ch, bars, nl, 0, len(t), t, te, nrows, ncols, 4,
                                        occupancy=occ)
    # one sampling event per registered contact, two units per event
    assert int(occ.sum()) == 2 * int(t.sum())
    assert int(band.sum()) + missed == int(t.sum())
    # units sit strictly inside (start + 1, end - 1): the last bin can only be partly reached
    assert occ[0] > 0 and occ[-1] <= occ.max()


def test_occupancy_draws_follow_the_contact_draws():
    # the occupancy events consume draws after the epoch's contacts: the first sampling epoch's
    # contacts are the same with and without tracking, later ones diverge
    rc, ch, bars, nl, t, te, nrows, ncols = setup(tiny(ncells=2))
    a, ma, sa = R.simulate_cells(rc, ch, bars, nl, 0, 2, t, te, nrows, ncols, 1)
    occ = np.zeros(_occ_len(ch, rc["bin_size"]), np.uint64)
    b, mb, sb = R.simulate_cells(rc, ch, bars, nl, 0, 2, t, te, nrows, ncols, 1, occupancy=occ)
    assert np.array_equal(sa[:, 1], sb[:, 1])  # burn-in is untouched
    assert not np.array_equal(a, b)
    assert int(a.sum()) + ma == int(b.sum()) + mb == int(t.sum())


def test_occupancy_matches_the_unit_marginals_of_loop_contacts():
    # loop contacts without noise register (bin(rev), bin(fwd)) of a sampled LEF; occupancy
    # registers bin(rev) and bin(fwd) of an independently sampled LEF of the same epoch. Both
    # estimate one profile: a dropped unit, a wrong bin origin or a shifted index shows up as
    # a mismatch of totals or a lag of the cross-correlation.
    wl = tiny(ncells=16, opts=dict(target_contact_density=20.0))
    c = OC.transform(dict(wl.options, sample_tad=False, noisify=False))
    rc = OC.to_rc(c)
    _, ch, bars, nl, t, te, nrows, ncols = setup(wl)
    import ctypes as C
    cfg = R.make_config(rc)
    t = np.zeros(16, np.uint64)
    te = R.lib().rc_plan_targets(C.byref(cfg), 16, nl, nrows * ncols,
                                 t.ctypes.data_as(C.POINTER(C.c_uint64)))
    occ = np.zeros(_occ_len(ch, rc["bin_size"]), np.uint64)
    band, missed, _ = R.simulate_cells(rc, ch, bars, nl, 0, 16, t, te, nrows, ncols, 8,
                                       occupancy=occ)
    m = band[: nrows * ncols].reshape(ncols, nrows).astype(np.float64)  # [col j][diag i]
    marg = np.zeros(ncols)
    for i in range(nrows):
        cols = np.arange(i, ncols)
        marg[cols] += m[cols, i]        # column bin j
        marg[cols - i] += m[cols, i]    # row bin j - i
    occ_f = occ[:ncols].astype(np.float64)
    assert abs(occ_f.sum() - marg.sum()) <= 2 * missed
    a = (occ_f - occ_f.mean()) / occ_f.std()
    b = (marg - marg.mean()) / marg.std()
    lags = {k: float(np.mean(a[max(0, k):len(a) + min(0, k)] * b[max(0, -k):len(b) - max(0, k)]))
            for k in range(-3, 4)}
    assert max(lags, key=lags.get) == 0
    assert lags[0] > 0.9


def test_state_log_has_one_record_per_completed_epoch():
    rc, ch, bars, nl, t, te, nrows, ncols = setup(tiny(ncells=4))
    band, missed, st, recs = R.simulate_cells(rc, ch, bars, nl, 0, 4, t, te, nrows, ncols, 2,
                                              state_log_cap=100_000)
    for cell in range(4):
        r = recs[recs["cell_id"] == cell]
        assert len(r) == int(st[cell, 0])
        assert np.array_equal(r["epoch"], np.arange(len(r)))
        # burn-in is a prefix; LEFs only get activated during it
        b = r["burnin"]
        k = int(b.sum())
        assert np.all(b[:k] == 1) and np.all(b[k:] == 0)
        assert np.all(np.diff(r["num_active_lefs"].astype(np.int64)) >= 0)
        assert np.all(r["num_active_lefs"][k:] == nl)
    assert np.all(recs["stalls_both"] <= np.minimum(recs["stalls_rev"], recs["stalls_fwd"]))
    assert np.all(recs["lef_bar"] + recs["primary"] + recs["secondary"]
                  <= recs["stalls_rev"] + recs["stalls_fwd"])
    assert np.all(recs["active_barriers"] <= recs["num_barriers"])
    assert np.all(recs["num_barriers"] == len(bars.pos))
    assert np.all(recs["avg_loop_size"] >= 0)
    # logging reads state only: the band is the one of a run without the log
    b2, m2, _ = R.simulate_cells(rc, ch, bars, nl, 0, 4, t, te, nrows, ncols, 2)
    assert np.array_equal(band, b2) and missed == m2


def test_state_log_overflow_is_reported():
    rc, ch, bars, nl, t, te, nrows, ncols = setup(tiny(ncells=2))
    try:
        R.simulate_cells(rc, ch, bars, nl, 0, 2, t, te, nrows, ncols, 1, state_log_cap=5)
    except RuntimeError as e:
        assert "capacity 5" in str(e)
    else:
        raise AssertionError("overflow not reported")


def test_state_counts_on_reference_collision_vectors():
    # the reference's own expected collisions ([index, event] pairs of Collision<u32f>, events
    # COLLISION 0x10 | CHROM_BOUNDARY 0x08 / LEF_BAR 0x04 / PRIMARY 0x02 / SECONDARY 0x01, and
    # the bypass marks without the COLLISION bit) counted by dump_stats' rules
    kats = json.load(open(os.path.join(HERE, "golden", "reference_kat.json")))
    checked = 0
    for k in kats:
        arr = k["arrays"]
        if "rev_collisions_expected" not in arr or "lefs" not in arr:
            continue
        rev, fwd, _ = KR.lef_arrays(arr["lefs"]["values"])
        enc = lambda v: np.array([KR.encode(p) for p in v], dtype=np.uint64)  # noqa: E731
        rv, fv = arr["rev_collisions_expected"]["values"], arr["fwd_collisions_expected"]["values"]
        got = R.count_state(rev, fwd, enc(rv), enc(fv), np.array([1, 0, 1], np.uint8))
        ev_r = np.array([e for _, e in rv])
        ev_f = np.array([e for _, e in fv])
        assert got["stalls_rev"] == int(np.sum(ev_r >= 0x10))
        assert got["stalls_fwd"] == int(np.sum(ev_f >= 0x10))
        assert got["stalls_both"] == int(np.sum((ev_r >= 0x10) & (ev_f >= 0x10)))
        for name, ev in (("lef_bar", 0x14), ("primary", 0x12), ("secondary", 0x11)):
            assert got[name] == int(np.sum(ev_r == ev) + np.sum(ev_f == ev)), (k["name"], name)
        assert got["active_barriers"] == 2 and got["num_barriers"] == 3
        loops = (fwd - rev).astype(np.float64)  # unbound LEFs: both units at the sentinel
        assert abs(got["avg_loop_size"] - loops.mean()) <= 1e-9 * max(1.0, loops.mean())
        checked += 1
    assert checked >= 15
