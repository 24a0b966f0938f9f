"""Host-side mirror of libmodle's simulate interface over the C-ABI.

`Simulation(config).run_simulate(genome)` plays the role of Simulation::run_simulate
(src/libmodle/cpu/scheduler_simulate.cpp:43-184): for every chromosome with barriers it plans
the LEF count and per-cell contact targets, simulates all cells on the GPU and returns the
chromosome's banded contact matrix in ContactMatrixDense<u32> layout
(data[col * nrows + (col - row)], nrows * ncols + 1 entries) with its missed-update count.
Option names follow modle::Config (simulation_config.hpp:24-135).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

_SET_BITS = {
    "rev_extrusion_speed": L.SET_REV_SPEED,
    "fwd_extrusion_speed": L.SET_FWD_SPEED,
    "extrusion_barrier_occupancy": L.SET_BARRIER_OCCUPANCY,
}


def _p(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def config_defaults() -> L.MgpuConfig:
    c = L.MgpuConfig()
    L.check(L.lib().mgpu_config_defaults(C.byref(c)))
    return c


def make_params(options: dict | None = None) -> L.MgpuParams:
    """Config defaults, the user's options, then Cli::transform_args (cli.cpp:1192-1215)."""
    c = config_defaults()
    mask = 0
    for k, v in (options or {}).items():
        if k == "num_cells":
            c.num_cells = int(v)
            continue
        if not hasattr(c, k):
            raise KeyError(f"unknown simulate option {k!r}")
        ftype = dict(L.MgpuConfig._fields_)[k]
        setattr(c, k, ftype(v).value if ftype in (C.c_int32,) else v)
        mask |= _SET_BITS.get(k, 0)
    p = L.MgpuParams()
    L.check(L.lib().mgpu_config_transform(C.byref(c), mask, C.byref(p)))
    return p


@dataclass
class Barriers:
    pos: np.ndarray
    dir: np.ndarray
    stp_active: np.ndarray
    stp_inactive: np.ndarray
    view: L.MgpuBarriers = field(init=False)

    def __post_init__(self):
        self.pos = np.ascontiguousarray(self.pos, dtype=np.uint64)
        self.dir = np.ascontiguousarray(self.dir, dtype=np.uint8)
        self.stp_active = np.ascontiguousarray(self.stp_active, dtype=np.float64)
        self.stp_inactive = np.ascontiguousarray(self.stp_inactive, dtype=np.float64)
        self.view = L.MgpuBarriers(len(self.pos), _p(self.pos, C.c_uint64), _p(self.dir, C.c_uint8),
                                   _p(self.stp_active, C.c_double),
                                   _p(self.stp_inactive, C.c_double))

    def __len__(self):
        return len(self.pos)


def chrom_view(name: str, size: int, start: int = 0, end: int | None = None) -> L.MgpuChrom:
    b = name.encode()
    ch = L.MgpuChrom(b, len(b), size, start, size if end is None else end)
    ch._name_bytes = b
    return ch


@dataclass
class ImportedBarriers:
    """A chromosome's barriers in the order of its interval tree (mgpu_import_barriers), with
    the BED interval of each."""
    bed_start: np.ndarray
    bed_end: np.ndarray
    pos: np.ndarray
    dir: np.ndarray
    stp_active: np.ndarray
    stp_inactive: np.ndarray

    def __len__(self):
        return len(self.pos)

    def subset(self, i0: int, i1: int) -> "ImportedBarriers":
        return ImportedBarriers(self.bed_start[i0:i1].copy(), self.bed_end[i0:i1].copy(),
                                self.pos[i0:i1].copy(), self.dir[i0:i1].copy(),
                                self.stp_active[i0:i1].copy(), self.stp_inactive[i0:i1].copy())


def import_barriers(params: L.MgpuParams, records, start: int, end: int,
                    name_puu=None) -> ImportedBarriers:
    """One chromosome's BED records (file order) as Genome::import_barriers keeps them
    (mgpu_import_barriers)."""
    n = len(records)
    recs = (L.MgpuBedRecord * max(n, 1))()
    for i, r in enumerate(records):
        recs[i] = L.MgpuBedRecord(r.start, r.end, r.score, r.strand.encode())
    m = max(n, 1)
    bs, be, pos = np.zeros(m, np.uint64), np.zeros(m, np.uint64), np.zeros(m, np.uint64)
    dirs, sa, si = np.zeros(m, np.uint8), np.zeros(m, np.float64), np.zeros(m, np.float64)
    puu = None
    if name_puu is not None:
        if len(name_puu) != n:
            raise ValueError("name_puu needs one value per record")
        puu = np.ascontiguousarray(name_puu, np.float64)
    nout = C.c_uint64(0)
    L.check(L.lib().mgpu_import_barriers(
        C.byref(params), recs, None if puu is None else _p(puu, C.c_double), n, start, end,
        _p(bs, C.c_uint64), _p(be, C.c_uint64), _p(pos, C.c_uint64), _p(dirs, C.c_uint8),
        _p(sa, C.c_double), _p(si, C.c_double), C.byref(nout)))
    k = nout.value
    return ImportedBarriers(bs[:k], be[:k], pos[:k], dirs[:k], sa[:k], si[:k])


def sort_barriers(imp: ImportedBarriers) -> Barriers:
    """The ExtrusionBarriers constructor's position sort (mgpu_sort_barriers) of a copy."""
    b = Barriers(imp.pos.copy(), imp.dir.copy(), imp.stp_active.copy(), imp.stp_inactive.copy())
    L.check(L.lib().mgpu_sort_barriers(len(b), _p(b.pos, C.c_uint64), _p(b.dir, C.c_uint8),
                                       _p(b.stp_active, C.c_double),
                                       _p(b.stp_inactive, C.c_double)))
    return b


def prepare_barriers(params: L.MgpuParams, records, start: int, end: int,
                     name_puu=None) -> Barriers:
    """Barrier tables of one chromosome's BED records (mgpu_prepare_barriers). name_puu, when
    given, holds each record's name field as a number: the record's not-bound stp, as the simulate
    CLI's --interpret-extrusion-barrier-name-as-not-bound-stp reads it."""
    n = len(records)
    recs = (L.MgpuBedRecord * max(n, 1))()
    for i, r in enumerate(records):
        recs[i] = L.MgpuBedRecord(r.start, r.end, r.score, r.strand.encode())
    pos = np.zeros(max(n, 1), np.uint64)
    dirs = np.zeros(max(n, 1), np.uint8)
    sa = np.zeros(max(n, 1), np.float64)
    si = np.zeros(max(n, 1), np.float64)
    nout = C.c_uint64(0)
    puu = None
    if name_puu is not None:
        if len(name_puu) != n:
            raise ValueError("name_puu needs one value per record")
        puu = np.ascontiguousarray(name_puu, np.float64)
    L.check(L.lib().mgpu_prepare_barriers_name_puu(
        C.byref(params), recs, None if puu is None else _p(puu, C.c_double), n, start, end,
        _p(pos, C.c_uint64), _p(dirs, C.c_uint8), _p(sa, C.c_double), _p(si, C.c_double),
        C.byref(nout)))
    k = nout.value
    return Barriers(pos[:k], dirs[:k], sa[:k], si[:k])


@dataclass
class ChromPlan:
    nlefs: int
    nrows: int
    ncols: int
    targets: np.ndarray
    target_epochs: int


def plan_chrom(params: L.MgpuParams, chrom: L.MgpuChrom, ncells: int) -> ChromPlan:
    nl, nr, nc, te = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
    targets = np.zeros(ncells, np.uint64)
    L.check(L.lib().mgpu_plan_chrom(C.byref(params), C.byref(chrom), ncells, C.byref(nl),
                                    C.byref(nr), C.byref(nc), _p(targets, C.c_uint64), C.byref(te)))
    return ChromPlan(nl.value, nr.value, nc.value, targets, te.value)


class Context:
    """One device context (a HIP stream and its timing events)."""

    def __init__(self, device: int = 0):
        self.ptr = C.c_void_p()
        L.check(L.lib().mgpu_init(device, C.byref(self.ptr)))

    def close(self):
        if self.ptr:
            L.lib().mgpu_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def phase_times(self, enable: bool = True, deterministic: bool = False):
        """Reads and clears the per-phase tick totals (100 MHz) and sets the timers: fast mode's
        16 phases, or with `deterministic` the wave-parallel kernel's 12 (L.DET_PHASE_NAMES)."""
        out = np.zeros(28, np.uint64)
        self.check(L.lib().mgpu_debug_phase_times(self.ptr, 1 if enable else 0,
                                                  _p(out, C.c_uint64), 28))
        return out[16:] if deterministic else out[:16]

    def check(self, rc):
        L.check(rc, self.ptr)


class Job:
    """Device-resident batch of chromosomes (mgpu_job_*): inputs uploaded once, bands in HBM."""

    def __init__(self, ctx: Context, params, chroms, barriers, plans, mode):
        """params: one MgpuParams for every chromosome, or a list with one per chromosome
        (mgpu_job_create_multi: they may differ in LEF density and processivity only)."""
        self.ctx = ctx
        n = len(chroms)
        self.plans = plans
        self._chroms = (L.MgpuChrom * n)(*chroms)
        self._bars = (L.MgpuBarriers * n)(*[b.view for b in barriers])
        self._keep = (chroms, barriers)
        self._nlefs = np.array([p.nlefs for p in plans], np.uint64)
        self._ncells = np.array([len(p.targets) for p in plans], np.uint64)
        self._targets = [np.ascontiguousarray(p.targets, np.uint64) for p in plans]
        self._tptrs = (L.U64P * n)(*[_p(t, C.c_uint64) for t in self._targets])
        self._tepochs = np.array([p.target_epochs for p in plans], np.uint64)
        self.ptr = C.c_void_p()
        if isinstance(params, (list, tuple)):
            if len(params) != n:
                raise ValueError("one parameter block per chromosome expected")
            self._params = (L.MgpuParams * n)(*params)
            create, pref = L.lib().mgpu_job_create_multi, self._params
        else:
            create, pref = L.lib().mgpu_job_create, C.byref(params)
        ctx.check(create(ctx.ptr, pref, n, self._chroms, self._bars,
                         _p(self._nlefs, C.c_uint64), _p(self._ncells, C.c_uint64),
                         self._tptrs, _p(self._tepochs, C.c_uint64), mode, C.byref(self.ptr)))

    def run(self) -> L.MgpuStats:
        st = L.MgpuStats()
        self.ctx.check(L.lib().mgpu_job_run(self.ptr, C.byref(st)))
        return st

    def block_sums(self, chrom, row, col, block_size: int) -> np.ndarray:
        """unsafe_get_block of the bands on the device (mgpu_job_block_sums): one u32 sum per
        (chromosome or task, row bin, column bin)."""
        c = np.ascontiguousarray(chrom, np.uint64)
        r = np.ascontiguousarray(row, np.uint64)
        q = np.ascontiguousarray(col, np.uint64)
        if not len(c) == len(r) == len(q):
            raise ValueError("chrom, row and col differ in length")
        out = np.zeros(max(len(c), 1), np.uint32)
        self.ctx.check(L.lib().mgpu_job_block_sums(self.ptr, len(c), _p(c, C.c_uint64),
                                                   _p(r, C.c_uint64), _p(q, C.c_uint64),
                                                   block_size, _p(out, C.c_uint32)))
        return out[:len(c)]

    def set_tuning(self, threads_per_cell: int = 0, global_working_set: bool | None = None,
                   tier_threads=None):
        """threads_per_cell 0: one launch per size tier (tier_threads overrides their sizes);
        otherwise one launch of that size over every cell."""
        t = self.tuning()
        t.threads_per_cell = threads_per_cell
        for k in range(5):
            t.tier_threads[k] = int(tier_threads[k]) if tier_threads and k < len(tier_threads) else 0
        if global_working_set is not None:
            t.global_working_set = 1 if global_working_set else 0
        self.ctx.check(L.lib().mgpu_job_set_tuning(self.ptr, C.byref(t)))

    def launches(self):
        """[(threads per cell, grid)] of the launches the next run makes."""
        t = self.tuning()
        return [(int(t.tier_threads[k]), int(t.tier_grid[k])) for k in range(int(t.launches))]

    def capture(self):
        """Records a run into a HIP graph that later runs replay."""
        self.ctx.check(L.lib().mgpu_job_capture(self.ptr))

    def release_graph(self):
        self.ctx.check(L.lib().mgpu_job_release_graph(self.ptr))

    def autotune(self, sample_cells: int = 0):
        """Times each workgroup size on a sample of cells, per size tier, and keeps the fastest;
        returns (the workgroup size of each launch, {threads: ms summed over the launches})."""
        t = L.MgpuTuning()
        ms = (C.c_float * 5)()
        self.ctx.check(L.lib().mgpu_job_autotune(self.ptr, sample_cells, C.byref(t), ms))
        sizes = [int(t.tier_threads[k]) for k in range(int(t.launches))]
        return sizes, dict(zip((64, 128, 256, 512, 1024), list(ms)))

    def tuning(self) -> L.MgpuTuning:
        t = L.MgpuTuning()
        self.ctx.check(L.lib().mgpu_job_get_tuning(self.ptr, C.byref(t)))
        return t

    def launch_stats(self):
        """[MgpuLaunchStats] of the last uncaptured run, one per launch."""
        out = (L.MgpuLaunchStats * 5)()
        n = C.c_uint32()
        self.ctx.check(L.lib().mgpu_job_launch_stats(self.ptr, out, 5, C.byref(n)))
        return [out[k] for k in range(min(n.value, 5))]

    def occupancy(self, c: int) -> np.ndarray:
        """Chromosome c's 1D LEF occupancy of the last run (u64 per bin); the job must have been
        created with track_1d_lef_position."""
        nb = C.c_uint64()
        self.ctx.check(L.lib().mgpu_job_read_occupancy(self.ptr, c, None, C.byref(nb)))
        out = np.zeros(nb.value, np.uint64)
        self.ctx.check(L.lib().mgpu_job_read_occupancy(self.ptr, c, _p(out, C.c_uint64), None))
        return out

    def set_state_log(self, capacity: int):
        """Room for `capacity` per-epoch state records in later runs (0: logging off)."""
        self.ctx.check(L.lib().mgpu_job_set_state_log(self.ptr, int(capacity)))

    def state_log(self):
        """The last run's state records, sorted by (chromosome, cell, epoch), as a structured
        array; None when the log overflowed (the second item is then the records needed)."""
        n = C.c_uint64()
        rc = L.lib().mgpu_job_read_state_log(self.ptr, None, 0, C.byref(n))
        if rc == L.MGPU_ERANGE and n.value:
            buf = (L.MgpuStateRecord * n.value)()
            rc = L.lib().mgpu_job_read_state_log(self.ptr, buf, n.value, C.byref(n))
            if rc == L.MGPU_ERANGE:
                return None, n.value
            self.ctx.check(rc)
            return np.frombuffer(buf, dtype=STATE_DTYPE, count=n.value).copy(), n.value
        self.ctx.check(rc)
        return np.zeros(0, STATE_DTYPE), 0

    def cell_stats(self) -> np.ndarray:
        """[cells, 4] u64 of the last run, in job order: epochs, burn-in epochs, contacts, and
        the engine words consumed (deterministic mode; 0 in fast mode)."""
        n = C.c_uint64(0)
        self.ctx.check(L.lib().mgpu_job_cell_stats(self.ptr, None, 0, C.byref(n)))
        out = np.zeros((n.value, 4), np.uint64)
        self.ctx.check(L.lib().mgpu_job_cell_stats(
            self.ptr, out.ctypes.data_as(C.POINTER(C.c_uint64)), out.size, C.byref(n)))
        return out

    def kernel(self) -> str:
        """Name of the simulation kernel this job launches (as rocprof lists it)."""
        buf = C.create_string_buffer(64)
        self.ctx.check(L.lib().mgpu_job_kernel(self.ptr, buf, 64))
        return buf.value.decode()

    def band(self, c: int):
        p = self.plans[c]
        out = np.zeros(p.nrows * p.ncols + 1, np.uint32)
        missed = C.c_uint64()
        self.ctx.check(L.lib().mgpu_job_read_band(self.ptr, c, _p(out, C.c_uint32), C.byref(missed)))
        return out, missed.value

    def close(self):
        if self.ptr:
            L.lib().mgpu_job_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Simulation:
    """Simulation::run_simulate over a workload (chromosomes + BED barrier records)."""

    def __init__(self, options: dict, ctx: Context | None = None):
        self.options = dict(options)
        self.params = make_params({k: v for k, v in options.items()
                                   if k not in ("num_cells", "nthreads")})
        self.num_cells = int(options.get("num_cells", 512))
        self.nthreads = int(options.get("nthreads", 0)) or (os.cpu_count() or 1)
        self.ctx = ctx or Context(0)
        self.lef_1d_occupancy = {}
        self.state_log = None
        self.state_log_tsv = ""

    def prepare(self, workload, simulate_chromosomes_wo_barriers: bool = False):
        chroms, bars, plans, names = [], [], [], []
        for name, size in workload.chroms:
            ch = chrom_view(name, size)
            b = prepare_barriers(self.params, workload.barriers_of(name), 0, size)
            if len(b) == 0 and not simulate_chromosomes_wo_barriers:
                continue  # scheduler_simulate.cpp:115-120
            chroms.append(ch)
            bars.append(b)
            plans.append(plan_chrom(self.params, ch, self.num_cells))
            names.append(name)
        return names, chroms, bars, plans

    def run_simulate(self, workload, mode: int = L.MODE_DETERMINISTIC, threads_per_cell: int = 0):
        """Returns ({chromosome: (band, missed, plan)}, stats). With track_1d_lef_position the
        chromosomes' 1D LEF occupancy lands in self.lef_1d_occupancy ({name: u64 bins}); with
        log_model_internal_state the per-epoch state records land in self.state_log and its
        text (the reference's TSV) in self.state_log_tsv."""
        names, chroms, bars, plans = self.prepare(workload)
        job = Job(self.ctx, self.params, chroms, bars, plans, mode)
        if threads_per_cell:
            job.set_tuning(threads_per_cell)
        stats = job.run()
        if self.params.log_model_internal_state:
            recs, need = job.state_log()
            if recs is None:  # the log overflowed: the run is deterministic, so size it and rerun
                job.set_state_log(need)
                stats = job.run()
                recs, need = job.state_log()
            self.state_log = recs
            self.state_log_tsv = format_state_log(recs, names, task_id_base(
                [len(p.targets) for p in plans], self.num_cells, self.nthreads))
        out = {}
        self.lef_1d_occupancy = {}
        for i, name in enumerate(names):
            band, missed = job.band(i)
            out[name] = (band, missed, plans[i])
            if self.params.track_1d_lef_position:
                self.lef_1d_occupancy[name] = job.occupancy(i)
        job.close()
        return out, stats


STATE_DTYPE = np.dtype([(n, {C.c_uint64: np.uint64, C.c_int32: np.int32, C.c_uint32: np.uint32,
                             C.c_double: np.float64}[t]) for n, t in L.MgpuStateRecord._fields_])

# Config::model_internal_state_log_header (simulation_config.hpp:126-131)
STATE_LOG_HEADER = ("task_id\tepoch\tcell_id\tchrom\tburnin\tbarrier_occupancy\tnum_active_lefs\t"
                    "num_stalls_rev\tnum_stalls_fwd\tnum_stalls_both\tnum_lef_bar_collisions\t"
                    "num_primary_lef_lef_collisions\tnum_secondary_lef_lef_collisions\t"
                    "avg_loop_size\n")


def task_id_base(cells_per_chrom, num_cells: int, nthreads: int):
    """First task id of each simulated chromosome as run_simulate numbers them
    (scheduler_simulate.cpp:70-78, 141-153): tasks are generated in batches of
    2 * min(16, num_cells / nthreads) (1 when num_cells <= nthreads) and ids advance over whole
    batches, including the unused tail of a chromosome's last batch."""
    deq = 1 if num_cells <= nthreads else min(16, num_cells // nthreads)
    enq = 2 * deq
    base, t = [], 0
    for n in cells_per_chrom:
        base.append(t)
        t += ((n + enq - 1) // enq) * enq
    return base


def format_state_log(recs, names, task_base) -> str:
    """The model internal-state log as Simulation::dump_stats writes it (simulation.cpp:1206-1220),
    one line per record, after the header."""
    lines = [STATE_LOG_HEADER]
    for r in recs:
        c = int(r["chrom"])
        nb = int(r["num_barriers"])
        occ = float(r["active_barriers"]) / nb if nb else float("nan")
        lines.append(f"{task_base[c] + int(r['cell_id'])}\t{int(r['epoch'])}\t{int(r['cell_id'])}\t"
                     f"{names[c]}\t{'True' if r['burnin'] else 'False'}\t{_fmt(occ)}\t"
                     f"{int(r['num_active_lefs'])}\t{int(r['stalls_rev'])}\t{int(r['stalls_fwd'])}\t"
                     f"{int(r['stalls_both'])}\t{int(r['lef_bar'])}\t{int(r['primary'])}\t"
                     f"{int(r['secondary'])}\t{_fmt(float(r['avg_loop_size']))}\n")
    return "".join(lines)


def _fmt(x: float) -> str:
    """fmt's default for a double: the shortest representation that round-trips."""
    if x != x:
        return "nan"
    if x == int(x) and abs(x) < 1e16:
        return str(int(x))
    return repr(x)


def occupancy_track(counts: np.ndarray) -> np.ndarray:
    """The values write_1d_lef_occupancy_to_disk hands to the BigWig writer
    (simulation.cpp:244-252): each bin's count over the chromosome's largest count, as float."""
    m = float(counts.max()) if len(counts) else 0.0
    return (counts.astype(np.float64) / m).astype(np.float32) if m else counts.astype(np.float32)


@dataclass
class WindowSpec:
    """One window task for WindowJob (mgpu_window_task): the window [start, end) of chromosome
    `name` of `size` bp, its sorted barriers (deletion applied), LEF count, cell id and targets."""
    name: str
    size: int
    start: int
    end: int
    barriers: Barriers
    nlefs: int
    cell_id: int
    target_contacts: int
    target_epochs: int


class WindowJob(Job):
    """Device-resident batch of perturbate / replay window tasks (mgpu_job_create_windows): task
    t is the job's chromosome t, with one cell and its own band over the window."""

    def __init__(self, ctx: Context, params, tasks, mode):
        self.ctx = ctx
        n = len(tasks)
        if n == 0:
            raise ValueError("no window tasks")
        bins, width = int(params.bin_size), int(params.diagonal_width)
        self.plans = []
        self._views = []
        arr = (L.MgpuWindowTask * n)()
        for k, t in enumerate(tasks):
            ncols = -(-(t.end - t.start) // bins)
            nrows = min(-(-width // bins), ncols)
            self.plans.append(ChromPlan(t.nlefs, nrows, ncols,
                                        np.array([t.target_contacts], np.uint64), t.target_epochs))
            ch = chrom_view(t.name, t.size, t.start, t.end)
            self._views.append(ch)
            arr[k] = L.MgpuWindowTask(ch, t.barriers.view, t.nlefs, t.cell_id, t.target_contacts,
                                      t.target_epochs)
        self._tasks = arr
        self._keep = tasks
        self.ptr = C.c_void_p()
        ctx.check(L.lib().mgpu_job_create_windows(ctx.ptr, C.byref(params), n, arr, mode,
                                                  C.byref(self.ptr)))


def band_block_sums(ctx: Context, band: np.ndarray, nrows: int, ncols: int, row, col,
                    block_size: int) -> np.ndarray:
    """unsafe_get_block on a band in host memory, on the device (mgpu_band_block_sums)."""
    b = np.ascontiguousarray(band, np.uint32)
    if len(b) != nrows * ncols + 1:
        raise ValueError("band length is not nrows * ncols + 1")
    r = np.ascontiguousarray(row, np.uint64)
    q = np.ascontiguousarray(col, np.uint64)
    out = np.zeros(max(len(r), 1), np.uint32)
    ctx.check(L.lib().mgpu_band_block_sums(ctx.ptr, _p(b, C.c_uint32), nrows, ncols, len(r),
                                           _p(r, C.c_uint64), _p(q, C.c_uint64), block_size,
                                           _p(out, C.c_uint32)))
    return out[:len(r)]
