"""The wave-parallel deterministic kernel's code (det_wave.h), run on the CPU with 64 emulated
lanes (tests/cpp/libdet_wave_emu.so), against the oracle: per-epoch state hashes, engine words
consumed, epochs and bands, bit for bit. The same header is what det_wave_kernel runs on the
GPU (checked there by tests/test_gpu_parity.py); this suite pins the parallel restatement itself
(speculative draws, scans, sorts, collision chains) on cases the oracle finishes in seconds."""
import ctypes as C
import os

import numpy as np
import pytest

from modlefork_amd import workloads as W
from oracle import config as OC
from oracle import refcpu as R

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "cpp", "libdet_wave_emu.so")

_lib = None
# path counters of det_wave.h (WDbg order), summed over every cell this module runs
DBG_NAMES = ["rank_serial", "repair_pass", "adjust_serial", "draw_redo", "contact_cut",
             "bypass", "chain_drop"]
DBG = np.zeros(len(DBG_NAMES), np.uint64)


def emu():
    global _lib
    if _lib is None:
        if not os.path.exists(EMU):
            pytest.skip("tests/cpp/libdet_wave_emu.so not built (run __graft_entry__.build())")
        R.lib()  # the emulator links the oracle library
        _lib = C.CDLL(EMU)
        _lib.emu_det_wave_cell.restype = C.c_int
        _lib.emu_det_wave_cell.argtypes = [
            C.POINTER(R.RcConfig), C.POINTER(R.RcChrom), C.POINTER(R.RcBarriers), C.c_uint64,
            C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint32), C.c_uint64,
            C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint64), C.c_uint64, C.POINTER(C.c_uint64),
            C.POINTER(C.c_uint64)]
        _lib.emu_det_wave_bytes.restype = C.c_uint64
        _lib.emu_det_wave_bytes.argtypes = [C.c_uint64] * 3
        _lib.emu_set_separate_cell.restype = None
        _lib.emu_set_separate_cell.argtypes = [C.c_int]
        _lib.emu_set_outputs.restype = None
        _lib.emu_set_outputs.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
    return _lib


def setup_case(size, nbar, seed, ncells, extra=None, bin_size=5000, width=1_000_000,
               clustered=False):
    rng = np.random.default_rng(seed)
    if clustered:
        centres = rng.choice(size - 200_000, max(1, nbar // 8), replace=False) + 100_000
        starts = np.unique(np.clip(
            (centres[:, None] + rng.integers(-3000, 3000, (len(centres), 8))).ravel(),
            0, size - 2))[:nbar]
    else:
        starts = np.sort(rng.choice(size - 1, nbar, replace=False))
    recs = [W.BedRecord("chrT", int(s), int(s) + 1, float(o), "+" if d else "-")
            for s, d, o in zip(starts, rng.random(len(starts)) < 0.5,
                               rng.uniform(0.5, 0.95, len(starts)))]
    opts = dict(bin_size=bin_size, diagonal_width=width, num_cells=ncells, seed=seed)
    if extra:
        opts.update(extra)
    c = OC.transform(opts)
    rc = OC.to_rc(c)
    bars = OC.prepare_barriers(recs, 0, size, c)
    nl = R.lib().rc_num_lefs(c["number_of_lefs_per_mbp"], size)
    ncols = R.lib().rc_band_ncols(c["bin_size"], size)
    nrows = R.lib().rc_band_nrows(c["diagonal_width"], c["bin_size"], size)
    t = np.zeros(ncells, np.uint64)
    cfg = R.make_config(rc)
    te = R.lib().rc_plan_targets(C.byref(cfg), ncells, nl, nrows * ncols,
                                 t.ctypes.data_as(C.POINTER(C.c_uint64)))
    ch = R.make_chrom("chrT", size, 0, size)
    return rc, ch, bars, nl, t, te, nrows, ncols


def run_emu(rc, ch, bars, nl, cell, target, te, nrows, ncols, max_trace=200_000):
    band = np.zeros(nrows * ncols + 1, np.uint32)
    missed = C.c_ulonglong(0)
    trace = np.zeros(max_trace, np.uint64)
    st = np.zeros(4, np.uint64)
    cfg = R.make_config(rc)
    r = emu().emu_det_wave_cell(C.byref(cfg), C.byref(ch), C.byref(bars.c), nl, cell, target, te,
                                band.ctypes.data_as(C.POINTER(C.c_uint32)), nrows,
                                C.byref(missed), trace.ctypes.data_as(C.POINTER(C.c_uint64)),
                                max_trace, st.ctypes.data_as(C.POINTER(C.c_uint64)),
                                DBG.ctypes.data_as(C.POINTER(C.c_uint64)))
    assert r == 0, f"emulated kernel failed (stats {st})"
    return band, int(missed.value), st, trace


def compare_cells(case, cells, max_trace=200_000):
    rc, ch, bars, nl, t, te, nrows, ncols = case
    for cell in cells:
        oband, omissed, ost, otrace = R.simulate_cell_trace(rc, ch, bars, nl, cell, int(t[cell]),
                                                            te, nrows, ncols, max_trace)
        band, missed, st, trace = run_emu(rc, ch, bars, nl, cell, int(t[cell]), te, nrows, ncols,
                                          max_trace)
        n = min(int(ost[0]), max_trace)
        diff = np.flatnonzero(trace[:n] != otrace[:n])
        assert len(diff) == 0, f"cell {cell}: first divergent epoch {diff[0]} of {n}; " \
                               f"oracle {ost} wave {tuple(int(x) for x in st)}"
        assert tuple(int(x) for x in ost) == tuple(int(x) for x in st), (ost, st)
        assert missed == omissed
        assert np.array_equal(band, oband), int(np.count_nonzero(band != oband))


def test_tiny_chromosome_cells_match_the_oracle():
    compare_cells(setup_case(2_000_000, 40, 3, 8), range(4))


def test_bypass_and_blocking_variants_match_the_oracle():
    # frequent secondary trials and bypasses, barriers that always / never block one side
    case = setup_case(3_000_000, 80, 7, 4, dict(probability_of_extrusion_unit_bypass=0.5,
                                                lef_bar_major_collision_pblock=0.9,
                                                lef_bar_minor_collision_pblock=0.3))
    compare_cells(case, range(2))
    case = setup_case(3_000_000, 80, 8, 4, dict(probability_of_extrusion_unit_bypass=0.0,
                                                tad_to_loop_contact_ratio=0.0))
    compare_cells(case, range(2))


def test_crowded_chromosome_with_units_stacked_at_its_ends():
    # few bp per LEF: units pile up at both ends, ties at equal positions, serial fall-backs
    case = setup_case(600_000, 30, 11, 4, dict(number_of_lefs_per_mbp=120,
                                               avg_lef_processivity=400_000),
                      bin_size=2000, width=200_000, clustered=True)
    compare_cells(case, range(3))


def test_many_new_lefs_with_tied_positions():
    """More than 64 LEFs bound in one epoch (several chunks of the new-LEF ranking) with equal
    binding positions among them, which must send the rank order to the serial sorters: 1 200
    LEFs on 60 kb with a 10 kb processivity rebind about 190 LEFs per epoch on 60 000
    positions."""
    before = int(DBG[DBG_NAMES.index("rank_serial")])
    case = setup_case(60_000, 10, 21, 2, dict(number_of_lefs_per_mbp=20_000,
                                              avg_lef_processivity=10_000),
                      bin_size=1000, width=20_000)
    compare_cells(case, range(1))
    assert int(DBG[DBG_NAMES.index("rank_serial")]) > before


def test_noise_free_tad_contacts_and_skipped_burnin():
    case = setup_case(2_000_000, 40, 5, 4, dict(tad_to_loop_contact_ratio=2.0, noisify_contacts=0,
                                                skip_burnin=True))
    compare_cells(case, range(2))


def test_small_configuration_cells_match_the_oracle():
    """20 Mbp, 5 kb bins, 3 Mbp band, 400 LEFs, 200 barriers (the small workload), 2 cells."""
    wl = W.small()
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
    ch = R.make_chrom(name, size, 0, size)
    compare_cells((rc, ch, bars, nl, t, te, nrows, ncols), (0, 37))


def test_cell_arrays_in_separate_memory_match_the_oracle():
    """The placement large cells use: ring, wave scratch and ziggurat tables in the block's LDS,
    the cell's arrays in a slot of their own (device memory on the GPU)."""
    emu().emu_set_separate_cell(1)
    try:
        compare_cells(setup_case(2_000_000, 40, 13, 4), range(2))
    finally:
        emu().emu_set_separate_cell(0)


# mgpu::StateRecord as the kernel writes it
EMU_RECORD = np.dtype([("item", np.uint32), ("epoch", np.uint32), ("burnin", np.uint32),
                       ("nact", np.uint32), ("active_barriers", np.uint32),
                       ("stalls_rev", np.uint32), ("stalls_fwd", np.uint32),
                       ("stalls_both", np.uint32), ("lef_bar", np.uint32), ("primary", np.uint32),
                       ("secondary", np.uint32), ("reserved", np.uint32),
                       ("avg_loop_size", np.float64)])


def test_occupancy_and_state_log_match_the_oracle():
    """With 1D LEF occupancy and the state log on (register_1d_lef_occupancy,
    register_contacts.cpp:200-232; dump_stats, simulation.cpp:1166-1224): the occupancy
    attempts draw after each epoch's contacts, so the band, the occupancy track and every
    per-epoch record must equal the oracle's run with the same outputs on."""
    rc, ch, bars, nl, t, te, nrows, ncols = setup_case(2_000_000, 40, 17, 4)
    cells = 2
    size = 2_000_000
    nbins = (size + 5000 - 1) // 5000
    occ = np.zeros(nbins, np.uint64)
    oband, omissed, ost, orecs = R.simulate_cells(rc, ch, bars, nl, 0, cells, t, te, nrows, ncols,
                                                  2, occupancy=occ, state_log_cap=100_000)
    docc = np.zeros(nbins, np.uint64)
    band = np.zeros(nrows * ncols + 1, np.uint32)
    cap = 20_000
    for cell in range(cells):
        rec = np.zeros(cap, EMU_RECORD)
        count = np.zeros(1, np.uint64)
        emu().emu_set_outputs(docc.ctypes.data, rec.ctypes.data, count.ctypes.data, cap)
        try:
            b, missed, st, _ = run_emu(rc, ch, bars, nl, cell, int(t[cell]), te, nrows, ncols,
                                       max_trace=1)
        finally:
            emu().emu_set_outputs(None, None, None, 0)
        band += b
        assert tuple(int(x) for x in st) == tuple(int(x) for x in ost[cell])
        n = int(count[0])
        assert n == int(st[0]) and n <= cap
        o = orecs[orecs["cell_id"] == cell]
        d = rec[:n]
        assert len(o) == n
        assert np.array_equal(d["epoch"], o["epoch"].astype(np.uint32))
        assert np.array_equal(d["burnin"], o["burnin"].astype(np.uint32))
        assert np.array_equal(d["nact"], o["num_active_lefs"])
        for f in ("active_barriers", "stalls_rev", "stalls_fwd", "stalls_both", "lef_bar",
                  "primary", "secondary"):
            assert np.array_equal(d[f].astype(np.uint64), o[f]), f
        assert np.array_equal(d["avg_loop_size"], o["avg_loop_size"])
    assert np.array_equal(docc, occ), int(np.count_nonzero(docc != occ))
    assert int(docc.sum()) == 2 * int(ost[:, 2].sum())
    assert np.array_equal(band, oband)


def test_every_exact_fallback_path_was_taken():
    """The cases above must reach each rare path of the parallel restatement at least once (they
    run first, in file order): serial rank order, odd-even repairs, the serial move walk, draws
    re-evaluated after a rejection, contact batches cut, bypasses and dropped chain members."""
    got = dict(zip(DBG_NAMES, (int(x) for x in DBG)))
    print(got)
    missing = [k for k, v in got.items() if v == 0]
    assert not missing, got
