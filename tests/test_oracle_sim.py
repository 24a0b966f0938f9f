"""Invariants of the oracle's whole-cell simulation and of its sorters."""
import ctypes as C

import numpy as np
import pytest

from modlefork_amd import workloads as W
from oracle import config as OC
from oracle import refcpu as R


def tiny(ncells=8, seed=5, size=2_000_000, opts=None):
    rng = np.random.default_rng(seed)
    starts = np.sort(rng.choice(size - 1, 40, replace=False))
    recs = [W.BedRecord("chrT", int(s), int(s) + 1, float(o), "+" if d else "-")
            for s, d, o in zip(starts, rng.random(40) < 0.5, rng.uniform(0.5, 0.95, 40))]
    o = dict(bin_size=5000, diagonal_width=1_000_000, num_cells=ncells, seed=seed)
    o.update(opts or {})
    return W.Workload("tiny", [("chrT", size)], recs, o)


def setup(wl):
    name, size = wl.chroms[0]
    c = OC.transform(wl.options)
    rc = OC.to_rc(c)
    bars = OC.prepare_barriers(wl.barriers_of(name), 0, size, c)
    nl = R.lib().rc_num_lefs(c["number_of_lefs_per_mbp"], size)
    ncols = R.lib().rc_band_ncols(c["bin_size"], size)
    nrows = R.lib().rc_band_nrows(c["diagonal_width"], c["bin_size"], size)
    ncells = wl.options["num_cells"]
    t = np.zeros(ncells, np.uint64)
    cfg = R.make_config(rc)
    te = R.lib().rc_plan_targets(C.byref(cfg), ncells, nl, nrows * ncols,
                                 t.ctypes.data_as(C.POINTER(C.c_uint64)))
    return rc, R.make_chrom(name, size, 0, size), bars, nl, t, te, nrows, ncols


def test_contacts_are_conserved_and_banded():
    rc, ch, bars, nl, t, te, nrows, ncols = setup(tiny())
    band, missed, st = R.simulate_cells(rc, ch, bars, nl, 0, len(t), t, te, nrows, ncols, 4)
    assert int(band.sum()) + missed == int(t.sum())
    assert np.all(st[:, 2] == t)
    assert band[-1] == 0  # the padding element stays untouched
    # pixels beyond the last column of each diagonal do not exist
    m = band[: nrows * ncols].reshape(ncols, nrows)
    for d in range(1, nrows):
        assert m[:d, d].sum() == 0


def test_cells_are_independent_of_threads_and_order():
    rc, ch, bars, nl, t, te, nrows, ncols = setup(tiny())
    a, ma, sa = R.simulate_cells(rc, ch, bars, nl, 0, 8, t, te, nrows, ncols, 1)
    b, mb, _ = R.simulate_cells(rc, ch, bars, nl, 0, 8, t, te, nrows, ncols, 8)
    parts = np.zeros_like(a)
    for lo in (0, 4):
        p, _, _ = R.simulate_cells(rc, ch, bars, nl, lo, lo + 4, t[lo:lo + 4], te, nrows, ncols, 2)
        parts += p
    assert np.array_equal(a, b) and ma == mb
    assert np.array_equal(a, parts)


def test_seed_changes_the_trajectory():
    r1 = setup(tiny(seed=5))
    wl = tiny(seed=5)
    wl.options["seed"] = 6
    r2 = setup(wl)
    a, _, _ = R.simulate_cells(*r1[:3], r1[3], 0, 2, r1[4], r1[5], r1[6], r1[7], 2)
    b, _, _ = R.simulate_cells(*r2[:3], r2[3], 0, 2, r2[4], r2[5], r2[6], r2[7], 2)
    assert not np.array_equal(a, b)


def test_loop_only_sampling_without_noise_lands_on_unit_pairs():
    # loop contacts without noise pair the two units of one LEF: both bins are integer positions
    # of bound units, so every contact lies within one LEF's processivity span
    wl = tiny(opts=dict(sample_tad_contacts=0, noisify=0))
    wl.options.pop("sample_tad_contacts")
    wl.options.pop("noisify")
    c = OC.transform(dict(wl.options, sample_tad=False, noisify=False))
    rc = OC.to_rc(c)
    assert rc["tad_to_loop_ratio"] == 0.0 and rc["noisify"] == 0
    _, ch, bars, nl, t, te, nrows, ncols = setup(wl)
    band, missed, _ = R.simulate_cells(rc, ch, bars, nl, 0, 4, t[:4], te, nrows, ncols, 4)
    assert int(band.sum()) + missed == int(t[:4].sum())


def test_skip_burnin_and_epoch_stopping():
    wl = tiny(ncells=2)
    c = OC.transform(dict(wl.options, skip_burnin=True))
    rc = OC.to_rc(c)
    _, ch, bars, nl, t, te, nrows, ncols = setup(wl)
    band, missed, st = R.simulate_cells(rc, ch, bars, nl, 0, 2, t[:2], te, nrows, ncols, 2)
    assert np.all(st[:, 1] == 0)
    c2 = OC.transform(dict(wl.options, stop_on_epochs=True, target_simulation_epochs=50))
    rc2 = OC.to_rc(c2)
    assert rc2["target_contact_density"] == -1.0


@pytest.mark.parametrize("n", [0, 1, 2, 23, 24, 25, 129, 1000, 5000])
def test_sorters_sort_and_permute(n):
    rng = np.random.default_rng(n)
    for dup in (False, True):
        key = rng.integers(0, 50 if dup else 10**9, max(n, 1)).astype(np.uint64)
        for fn in (R.lib().rc_pdq_sort_idx, R.lib().rc_split_sort_idx):
            idx = rng.permutation(n).astype(np.uint64)
            before = idx.copy()
            fn(idx.ctypes.data_as(C.POINTER(C.c_size_t)), n,
               key.ctypes.data_as(C.POINTER(C.c_uint64)))
            assert sorted(idx.tolist()) == sorted(before.tolist())
            assert np.all(np.diff(key[idx.astype(np.int64)].astype(np.int64)) >= 0)


def test_split_sort_of_nearly_sorted_input_keeps_equal_runs_stable():
    # a sorted run with a few displaced elements: the untouched elements keep their relative
    # order among equal keys (stable merge of the retained subsequence)
    key = np.repeat(np.arange(100, dtype=np.uint64), 3)
    idx = np.arange(len(key), dtype=np.uint64)
    R.lib().rc_split_sort_idx(idx.ctypes.data_as(C.POINTER(C.c_size_t)), len(idx),
                              key.ctypes.data_as(C.POINTER(C.c_uint64)))
    assert idx.tolist() == list(range(len(key)))


def test_rank_lefs_breaks_ties_by_binding_epoch():
    rev = np.array([10, 5, 10, 10, 3], np.uint64)
    fwd = np.array([20, 9, 20, 20, 4], np.uint64)
    ep = np.array([7, 1, 3, 5, 2], np.uint64)
    rr = np.zeros(5, np.uint64)
    fr = np.zeros(5, np.uint64)
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint64))
    R.lib().rc_rank_lefs(p(rev), p(fwd), p(ep), 5, rr.ctypes.data_as(C.POINTER(C.c_size_t)),
                         fr.ctypes.data_as(C.POINTER(C.c_size_t)), 0, 1)
    assert rr.tolist() == [4, 1, 2, 3, 0]  # rev ties: ascending binding epoch
    assert fr.tolist() == [4, 1, 0, 3, 2]  # fwd ties: descending binding epoch
