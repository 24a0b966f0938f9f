"""Deterministic mode on the GPU against the CPU oracle: the reference's known-answer step
tests, the samplers draw for draw, and whole cells / whole chromosomes bit for bit.

Every test calls the HIP code through the C-ABI (libmodle_gpu.so)."""
import ctypes as C
import sys

import numpy as np
import pytest

import kat_replay
from device_backend import DeviceBackend
from modlefork_amd import _lib as L
from modlefork_amd import simulation as S
from modlefork_amd import workloads as W
from oracle import config as OC
from oracle import refcpu as R

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


KAT = kat_replay.kat_cases()


@pytest.mark.parametrize("case", KAT, ids=[c["name"][:40] for c in KAT])
def test_device_steps_match_reference_kat(ctx, case):
    say(f"KAT {case['name']}")
    bad = kat_replay.replay(case, DeviceBackend(ctx))
    assert not bad, bad


SAMPLERS = [
    (0, 0, 0, 0), (1, 4000.0, 200.0, 0), (1, 0.0, 1.0, 0), (2, 0.75, 0, 0), (2, 0.0266667, 0, 0),
    (3, 0, 0, 0), (4, 0, 19_999_999, 0), (4, 0, 399, 0), (4, 1000, 1000, 0), (5, 2.13903743, 0, 0),
    (5, 26.7379679, 0, 0), (5, 150.0, 0, 0), (6, 64, 1 / 6, 0), (6, 320, 1 / 6, 0),
    (6, 40, 0.9, 0), (8, 3.4426198558966521, 0, 0), (7, 0.0, 5000.0, 0.001),
]


@pytest.mark.parametrize("kind,a,b,c", SAMPLERS)
def test_samplers_draw_for_draw(ctx, kind, a, b, c):
    n = 20000
    say(f"sampler kind={kind} a={a} b={b}")
    for seed in (0, 752741483, 10556020843759504871):
        want = R.draw_sequence(seed, kind, n, a, b, c)
        got = np.zeros(n, np.uint64)
        ctx.check(L.lib().mgpu_debug_draws(ctx.ptr, seed, kind, a, b, c, n,
                                           got.ctypes.data_as(C.POINTER(C.c_uint64))))
        if kind != 7:
            assert np.array_equal(got, want), (seed, np.flatnonzero(got != want)[:5])
        else:
            # GEV noise is mu + sigma (1 - t^xi) / xi with t = -log(u): the device and host libm
            # may round log/pow in the last bit, and 1 - t^xi cancels ~3 digits, so the noise
            # may differ by ~1e-9 bp. Positions are truncated to integers, which must agree.
            g = got.view(np.float64)
            w = want.view(np.float64)
            diff = np.abs(g - w)
            assert diff.max() < 1e-6, diff.max()
            for pos in (1_000_000.0, 12_345_678.0, 249_999_999.0):
                assert np.array_equal(np.floor(pos - g), np.floor(pos - w))
                assert np.array_equal(np.floor(pos + g), np.floor(pos + w))
            say(f"  gev seed={seed}: {np.count_nonzero(diff)} of {n} differ, max {diff.max():.2e}")


def small_setup(size=2_000_000, ncells=8, nbar=40, seed=3):
    rng = np.random.default_rng(seed)
    starts = np.sort(rng.choice(size - 1, nbar, replace=False))
    recs = [W.BedRecord("chrT", int(s), int(s) + 1, float(o), "+" if d else "-")
            for s, d, o in zip(starts, rng.random(nbar) < 0.5, rng.uniform(0.5, 0.95, nbar))]
    wl = W.Workload("tiny", [("chrT", size)], recs,
                    dict(bin_size=5000, diagonal_width=1_000_000, num_cells=ncells, seed=seed))
    return wl


def oracle_chrom(wl, nthreads=8):
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
    band, missed, st = R.simulate_cells(rc, ch, bars, nl, 0, ncells, t, te, nrows, ncols, nthreads)
    return band, missed, st, (rc, ch, bars, nl, t, te, nrows, ncols)


def first_divergence(ctx, wl, cell, setup):
    rc, ch, bars, nl, t, te, nrows, ncols = setup
    _, _, st, otrace = R.simulate_cell_trace(rc, ch, bars, nl, cell, int(t[cell]), te, nrows, ncols)
    name, size = wl.chroms[0]
    p = S.make_params({k: v for k, v in wl.options.items() if k != "num_cells"})
    dch = S.chrom_view(name, size)
    dbars = S.prepare_barriers(p, wl.barriers_of(name), 0, size)
    mx = len(otrace)
    dtrace = np.zeros(mx, np.uint64)
    out = np.zeros(4, np.uint64)
    ctx.check(L.lib().mgpu_debug_cell_trace(ctx.ptr, C.byref(p), C.byref(dch), C.byref(dbars.view),
                                            nl, cell, int(t[cell]), te, mx,
                                            dtrace.ctypes.data_as(C.POINTER(C.c_uint64)),
                                            out.ctypes.data_as(C.POINTER(C.c_uint64))))
    n = int(st[0])
    diff = np.flatnonzero(dtrace[:n] != otrace[:n])
    return (int(diff[0]) if len(diff) else None), st, tuple(int(x) for x in out)


def device_chrom(ctx, wl):
    sim = S.Simulation(wl.options, ctx)
    out, stats = sim.run_simulate(wl, L.MODE_DETERMINISTIC)
    name = wl.chroms[0][0]
    return out[name][0], out[name][1], stats


KERNELS = {"wave": 0, "lane": 1}  # mgpu_debug_det_kernel: wave-parallel / one-lane kernel


@pytest.fixture(params=list(KERNELS))
def det_kernel(ctx, request):
    ctx.check(L.lib().mgpu_debug_det_kernel(ctx.ptr, KERNELS[request.param]))
    yield request.param
    ctx.check(L.lib().mgpu_debug_det_kernel(ctx.ptr, 0))


def test_tiny_chromosome_cell_traces_match(ctx, det_kernel):
    wl = small_setup()
    say(f"tiny: oracle ({det_kernel} kernel)")
    band, missed, st, setup = oracle_chrom(wl)
    for cell in (0, 1):
        say(f"tiny: trace cell {cell}")
        div, ost, dst = first_divergence(ctx, wl, cell, setup)
        assert div is None, f"cell {cell}: first divergent epoch {div}; oracle {ost} device {dst}"
        assert tuple(int(x) for x in ost) == dst


def test_tiny_chromosome_band_bit_exact(ctx, det_kernel):
    wl = small_setup()
    band, missed, st, _ = oracle_chrom(wl)
    say(f"tiny: device ({det_kernel} kernel)")
    dband, dmissed, stats = device_chrom(ctx, wl)
    assert dmissed == missed
    assert int(stats.epochs) == int(st[:, 0].sum())
    assert int(stats.burnin_epochs) == int(st[:, 1].sum())
    assert np.array_equal(dband, band), int(np.count_nonzero(dband != band))


VARIANTS = {
    # frequent LEF-BAR and LEF-LEF trials (every trial draws), bypasses
    "trials": dict(probability_of_extrusion_unit_bypass=0.5, lef_bar_major_collision_pblock=0.9,
                   lef_bar_minor_collision_pblock=0.3),
    # no bypass draws, loop contacts only
    "no-bypass": dict(probability_of_extrusion_unit_bypass=0.0, tad_to_loop_contact_ratio=0.0),
    # crowded: units stacked at both chromosome ends, ties in position and binding epoch
    "crowded": dict(number_of_lefs_per_mbp=120, avg_lef_processivity=400_000),
    # TAD contacts without noise, burn-in skipped
    "tad-no-noise": dict(tad_to_loop_contact_ratio=2.0, noisify_contacts=0, skip_burnin=True),
}


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_parameter_variants_band_bit_exact(ctx, variant):
    """The wave-parallel kernel's rare paths (redrawn samples, serial rank orders and move walks,
    bypass chains) against the oracle on whole chromosomes, 8 cells each."""
    size = 600_000 if variant == "crowded" else 3_000_000
    wl = small_setup(size=size, ncells=8, nbar=30 if variant == "crowded" else 80, seed=7)
    wl.options.update(VARIANTS[variant])
    if variant == "crowded":
        wl.options.update(bin_size=2000, diagonal_width=200_000)
    band, missed, st, _ = oracle_chrom(wl)
    say(f"variant {variant}: device")
    dband, dmissed, stats = device_chrom(ctx, wl)
    assert dmissed == missed
    assert int(stats.epochs) == int(st[:, 0].sum())
    assert np.array_equal(dband, band), int(np.count_nonzero(dband != band))


def test_small_config_band_bit_exact(ctx, det_kernel):
    """20 Mbp, 5 kb bins, 3 Mbp band, 400 LEFs, 200 barriers, 64 cells, seed 0."""
    wl = W.small()
    say(f"small: oracle (64 cells, {det_kernel} kernel)")
    band, missed, st, _ = oracle_chrom(wl)
    say("small: device")
    dband, dmissed, stats = device_chrom(ctx, wl)
    say(f"small: device kernel {stats.kernel_ms:.1f} ms")
    assert int(band.sum()) + missed == 64 * 37500
    assert dmissed == missed
    assert int(stats.epochs) == int(st[:, 0].sum())
    assert np.array_equal(dband, band), int(np.count_nonzero(dband != band))


def test_cell_subset_adds_up(ctx):
    """Cells are independent: simulating [0,4) and [4,8) separately sums to [0,8)."""
    wl = small_setup(ncells=8)
    name, size = wl.chroms[0]
    p = S.make_params({k: v for k, v in wl.options.items() if k != "num_cells"})
    ch = S.chrom_view(name, size)
    bars = S.prepare_barriers(p, wl.barriers_of(name), 0, size)
    plan = S.plan_chrom(p, ch, 8)
    acc = np.zeros(plan.nrows * plan.ncols + 1, np.uint32)
    for begin in (0, 4):
        missed = C.c_uint64()
        tg = np.ascontiguousarray(plan.targets[begin:begin + 4])
        ctx.check(L.lib().mgpu_simulate_chrom(
            ctx.ptr, C.byref(p), C.byref(ch), C.byref(bars.view), plan.nlefs, begin, 4,
            tg.ctypes.data_as(C.POINTER(C.c_uint64)), plan.target_epochs, L.MODE_DETERMINISTIC,
            acc.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(missed), None))
    full, _, _ = device_chrom(ctx, wl)
    assert np.array_equal(acc, full)


UNIT_KAT = None


def _unit_kat():
    global UNIT_KAT
    if UNIT_KAT is None:
        import unit_kat_replay

        UNIT_KAT = unit_kat_replay.load()
    return UNIT_KAT


@pytest.mark.parametrize("mode", [L.MODE_DETERMINISTIC, L.MODE_FAST])
@pytest.mark.parametrize("case", ["CMatrix simple", "CMatrix in/decrement"])
def test_device_band_increments_match_reference_kat(ctx, mode, case):
    """contact_matrix_dense_test.cpp's increments applied by the device's band code (each mode's
    own), including the out-of-band increment at (11, 0) of a 10-row band that only counts as
    missed; a coordinate beyond the matrix is refused."""
    import unit_kat_replay as K

    def increment(band, missed, nrows, ncols, rows, cols):
        m = C.c_uint64(missed)
        r = np.asarray(rows, np.uint64)
        c = np.asarray(cols, np.uint64)
        ctx.check(L.lib().mgpu_debug_band_increments(
            ctx.ptr, mode, nrows, ncols, len(r), r.ctypes.data_as(L.U64P),
            c.ctypes.data_as(L.U64P), band.ctypes.data_as(L.U32P), C.byref(m)))
        return m.value

    say(f"band KAT {case}, mode {mode}")
    K.replay_band(_unit_kat()["cmatrix"][case], increment)
    band = np.zeros(10 * 20 + 1, np.uint32)
    with pytest.raises(L.MgpuError, match="outside the matrix"):
        increment(band, 0, 10, 20, [25], [25])


def test_device_collision_words_match_reference_kat(ctx):
    """collision_encoding_test.cpp's encodings and occurred / avoided checks on the device's
    collision words (deterministic mode's reference layout)."""
    import unit_kat_replay as K

    def words(idx, ev):
        i = np.array([idx], np.uint64)
        e = np.array([ev], np.uint8)
        w, ix = np.zeros(1, np.uint64), np.zeros(1, np.uint64)
        evo, fl = np.zeros(1, np.uint32), np.zeros(1, np.uint32)
        ctx.check(L.lib().mgpu_debug_collisions(
            ctx.ptr, 1, i.ctypes.data_as(L.U64P), e.ctypes.data_as(L.U8P), w.ctypes.data_as(L.U64P),
            ix.ctypes.data_as(L.U64P), evo.ctypes.data_as(L.U32P), fl.ctypes.data_as(L.U32P)))
        return int(w[0]), int(ix[0]), int(evo[0]), int(fl[0])

    assert K.replay_collisions(_unit_kat()["collisions"]["cases"], words) == 56


def test_cpp_caller_simulates_through_the_c_abi():
    """The C++ caller (tests/cpp/abi_caller.cpp, no ctypes) runs the integration's call sequence,
    mgpu_init -> mgpu_prepare_barriers -> mgpu_plan_chrom -> mgpu_simulate_chrom, on a 2 Mbp
    chromosome with 40 barriers and 4 cells: the deterministic band equals the oracle's bit for
    bit, with its missed count, and fast mode conserves every contact."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "abi_caller")
    assert os.path.exists(exe), "build the C++ caller first (__graft_entry__.build())"
    say("C++ caller: run")
    r = subprocess.run([exe, "run"], capture_output=True, text=True, timeout=240)
    say(r.stdout.strip())
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr
