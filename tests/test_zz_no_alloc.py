"""mgpu_simulate_chrom_noalloc (fast mode, no device allocation: its counters, barrier tables,
work items and statistics live in the library's per-GPU static scratch memory) against
mgpu_simulate_chrom on the same inputs, through the C-ABI: identical bands, missed counts and
statistics on two calls in a row, the second with a smaller input (fewer barriers, LEFs and
cells than the first left in the scratch memory), and a refusal of inputs beyond its limits."""
import ctypes as C

import numpy as np
import pytest

from modlefork_amd import _lib as L
from modlefork_amd import simulation as S
from modlefork_amd import workloads as W

pytestmark = pytest.mark.gpu


def _inputs(ctx, wl):
    sim = S.Simulation(wl.options, ctx)
    names, chroms, bars, plans = sim.prepare(wl)
    return sim.params, chroms[0], bars[0], plans[0]


def _allocating(ctx, p, ch, bars, plan, ncells):
    band = np.zeros(plan.nrows * plan.ncols + 1, np.uint32)
    missed = C.c_uint64(0)
    st = L.MgpuStats()
    t = np.ascontiguousarray(plan.targets[:ncells], np.uint64)
    ctx.check(L.lib().mgpu_simulate_chrom(
        ctx.ptr, C.byref(p), C.byref(ch), C.byref(bars.view), plan.nlefs, 0, ncells,
        t.ctypes.data_as(L.U64P), plan.target_epochs, L.MODE_FAST, band.ctypes.data_as(L.U32P),
        C.byref(missed), C.byref(st)))
    return band, missed.value, st


class DeviceBuffer:
    """Device memory owned by the test (the caller's band), through the HIP runtime the library
    itself links."""

    def __init__(self, nbytes):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.nbytes = nbytes
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(nbytes)) == 0
        assert self.hip.hipMemset(self.ptr, 0, C.c_size_t(nbytes)) == 0
        assert self.hip.hipDeviceSynchronize() == 0  # the library works on its own stream

    def read_u32(self):
        out = np.zeros(self.nbytes // 4, np.uint32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr,
                                  C.c_size_t(self.nbytes), 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def _noalloc(ctx, p, ch, bars, plan, ncells, cell_begin=0):
    dev = DeviceBuffer(4 * (plan.nrows * plan.ncols + 1))
    try:
        missed = C.c_uint64(0)
        st = L.MgpuStats()
        t = np.ascontiguousarray(plan.targets[cell_begin:cell_begin + ncells], np.uint64)
        rc = L.lib().mgpu_simulate_chrom_noalloc(
            ctx.ptr, C.byref(p), C.byref(ch), C.byref(bars.view), plan.nlefs, cell_begin, ncells,
            t.ctypes.data_as(L.U64P), plan.target_epochs, dev.ptr, C.byref(missed), C.byref(st))
        ctx.check(rc)
        return dev.read_u32(), missed.value, st
    finally:
        dev.free()


def _small_chromosome():
    rng = np.random.default_rng(21)
    size = 5_000_000
    recs = W._uniform_barriers(rng, "chrQ", size, 50, lambda n: rng.uniform(0.5, 0.95, n))
    return W.Workload("chrQ", [("chrQ", size)], recs,
                      dict(bin_size=5_000, diagonal_width=1_000_000, number_of_lefs_per_mbp=20.0,
                           num_cells=4, target_contact_density=2.0, seed=9))


def test_noalloc_matches_the_allocating_entry_point():
    ctx = S.Context(0)
    try:
        for wl, ncells in ((W.small(ncells=16), 16), (_small_chromosome(), 4)):
            p, ch, bars, plan = _inputs(ctx, wl)
            print(f"{wl.name}: {plan.nlefs} LEFs, {len(bars)} barriers, {ncells} cells",
                  flush=True)
            a_band, a_missed, a_st = _allocating(ctx, p, ch, bars, plan, ncells)
            b_band, b_missed, b_st = _noalloc(ctx, p, ch, bars, plan, ncells)
            assert np.array_equal(a_band, b_band)
            assert a_missed == b_missed
            assert (a_st.cells, a_st.epochs, a_st.burnin_epochs, a_st.contacts, a_st.dropped) == \
                (b_st.cells, b_st.epochs, b_st.burnin_epochs, b_st.contacts, b_st.dropped)
            assert int(b_band.sum()) + b_missed + b_st.dropped == int(plan.targets[:ncells].sum())
    finally:
        ctx.close()


def test_noalloc_refuses_inputs_beyond_its_limits():
    ctx = S.Context(0)
    try:
        p, ch, bars, plan = _inputs(ctx, W.small(ncells=16))
        too_many = np.zeros(16385, np.uint64)
        missed = C.c_uint64(0)
        rc = L.lib().mgpu_simulate_chrom_noalloc(
            ctx.ptr, C.byref(p), C.byref(ch), C.byref(bars.view), plan.nlefs, 0, 16385,
            too_many.ctypes.data_as(L.U64P), plan.target_epochs, C.c_void_p(0),
            C.byref(missed), None)
        assert rc == L.MGPU_EINVAL  # null band is refused before the cell-count check
        dev = DeviceBuffer(4 * (plan.nrows * plan.ncols + 1))
        try:
            rc = L.lib().mgpu_simulate_chrom_noalloc(
                ctx.ptr, C.byref(p), C.byref(ch), C.byref(bars.view), plan.nlefs, 0, 16385,
                too_many.ctypes.data_as(L.U64P), plan.target_epochs, dev.ptr, C.byref(missed),
                None)
            assert rc == L.MGPU_ERANGE
            assert b"16384" in L.lib().mgpu_last_error(ctx.ptr)
        finally:
            dev.free()
    finally:
        ctx.close()
