"""modle perturbate and modle replay on the device: windows of a chromosome simulated with
deletions that switch barriers off, every (window, deletion) task one cell of one persistent
launch, and the feature-pair readout of the bands.

Host side of the path, mirroring the reference's scheduler (the task plan of
Simulation::run_perturbate, src/libmodle/cpu/scheduler_perturbate.cpp:97-369; the window helpers
of scheduler_common.cpp:42-127; the readout of Simulation::simulate_window,
scheduler_perturbate.cpp:438-560; the task file of fmt::formatter<TaskPW>,
simulation_impl.hpp:162-170, and TaskPW::from_string, simulation.cpp:640-700; Simulation::
run_replay, scheduler_replay.cpp:39-154). The simulation and the block sums run in
libmodle_gpu (mgpu_job_create_windows, mgpu_job_block_sums, mgpu_band_block_sums); interval
trees are ordered with the library's pdq_sort (mgpu_pdq_order), as the reference's IITree is.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from . import simulation as S

USIZE_MAX = (1 << 64) - 1


@dataclass
class Feature:
    """A BED6 record of a feature track (--feature-beds)."""
    chrom: str
    start: int
    end: int
    name: str = ""
    score: float = 0.0
    strand: str = "."


def pdq_order(keys) -> np.ndarray:
    """Indices ordered by key as cpp-sort's pdq_sort leaves them (mgpu_pdq_order)."""
    k = np.ascontiguousarray(keys, np.uint64)
    out = np.zeros(max(len(k), 1), np.uint32)
    L.check(L.lib().mgpu_pdq_order(k.ctypes.data_as(L.U64P), len(k),
                                   out.ctypes.data_as(L.U32P)))
    return out[:len(k)]


class IntervalTree:
    """The reference's IITree over [start, end) intervals (interval_tree_impl.hpp): intervals
    ordered by start with pdq_sort (make_BST, :58-68); equal_range(start, end) returns the index
    range from the first to one past the last interval overlapping [start, end) (:139-218), or
    (n, n) when none does."""

    def __init__(self, starts, ends, data=None, indexed: bool = False):
        """indexed: the intervals are already in tree order (no further sort)."""
        order = np.arange(len(starts), dtype=np.uint32) if indexed else pdq_order(starts)
        self.start = np.asarray(starts, np.uint64)[order]
        self.end = np.asarray(ends, np.uint64)[order]
        self.order = order
        self.data = [data[i] for i in order] if data is not None else None

    def __len__(self):
        return len(self.start)

    def equal_range(self, start: int, end: int):
        hit = np.nonzero((self.start < end) & (self.end > start))[0]
        if len(hit) == 0:
            return len(self), len(self)
        return int(hit[0]), int(hit[-1]) + 1


def generate_deletions(barriers: S.ImportedBarriers, deletion_size: int) -> IntervalTree:
    """Simulation::generate_deletions for one chromosome (scheduler_common.cpp:51-68): one
    deletion per barrier, in the chromosome's barrier order, ending right after the barrier
    (shortened at the chromosome's 5' end), indexed as a BED tree."""
    starts, ends = [], []
    for p in barriers.pos.tolist():
        size = min(p + 1, deletion_size)
        b = p + 1 - size
        starts.append(b)
        ends.append(b + size)
    return IntervalTree(starts, ends)


def deletion_tree(intervals) -> IntervalTree:
    """A --deletion-list BED3 file's records of one chromosome (Simulation::import_deletions,
    scheduler_common.cpp:42-49): [(start, end)] in file order, indexed as a BED tree."""
    return IntervalTree([a for a, _ in intervals], [b for _, b in intervals])


def feature_tree(features) -> IntervalTree:
    """One feature track of one chromosome (Chromosome::get_features()[k]), in file order."""
    return IntervalTree([f.start for f in features], [f.end for f in features], list(features))


@dataclass
class Task:
    """Simulation::TaskPW (simulation.hpp:78-95) of one (window, deletion)."""
    id: int
    chrom: str
    chrom_size: int
    cell_id: int
    num_target_epochs: int
    num_target_contacts: int
    num_lefs: int
    deletion_begin: int
    deletion_size: int
    window_start: int
    window_end: int
    active_window_start: int
    active_window_end: int
    barrier_range: tuple  # (i0, i1) into the chromosome's barrier tree
    feats1_range: tuple
    feats2_range: tuple
    chrom_end: int = 0

    @property
    def num_barriers(self) -> int:
        return self.barrier_range[1] - self.barrier_range[0]


@dataclass
class ChromInput:
    """One chromosome of a perturbate run: its range, its barriers in interval-tree order
    (simulation.import_barriers), the BED intervals of those barriers and its two feature
    tracks."""
    name: str
    size: int
    start: int
    end: int
    barriers: S.ImportedBarriers
    feats1: list
    feats2: list
    barrier_tree: IntervalTree = field(init=False)
    feat_trees: tuple = field(init=False)

    def __post_init__(self):
        # import_barriers leaves the barriers in the chromosome's interval-tree order
        self.barrier_tree = IntervalTree(self.barriers.bed_start, self.barriers.bed_end,
                                         indexed=True)
        if np.any(np.diff(self.barrier_tree.start.astype(np.int64)) < 0):
            raise ValueError("barriers must be in the chromosome's interval-tree order")
        self.feat_trees = (feature_tree(self.feats1), feature_tree(self.feats2))


def windows(chrom_end: int, diagonal_width: int):
    """The outer and active windows run_perturbate visits on a chromosome
    (scheduler_perturbate.cpp:245-253, advance_window at scheduler_common.cpp:70-80): yields
    (window_start, window_end, active_window_start, active_window_end) before the task's clamp to
    the chromosome end; the first window's outer end is 4 D and its active end 3 D."""
    D = diagonal_width
    ws, we, aws, awe = 0, 4 * D, 0, 3 * D
    while True:
        yield ws, we, aws, awe
        aws += D
        awe = min(aws + 2 * D, chrom_end)
        ws = aws - D
        we = min(awe + D, chrom_end)
        if not aws < chrom_end:
            return


def plan(params: L.MgpuParams, chroms, deletions=None, deletion_size: int = 10_000,
         target_simulation_epochs: int = 0):
    """The task list of Simulation::run_perturbate (scheduler_perturbate.cpp:189-342).
    chroms: ChromInput in genome order. deletions: {chrom name: IntervalTree} from a deletion
    list, or None to generate one deletion per barrier (deletion_size bp). Chromosomes without
    barriers or without both feature tracks are skipped, as are windows without a barrier or
    without a feature of either track in the active window."""
    D = int(params.diagonal_width)
    bins = int(params.bin_size)
    density = float(params.target_contact_density)
    tasks = []
    task_id = 0
    for ch in chroms:
        if len(ch.barriers) == 0 or not ch.feats1 or not ch.feats2:
            continue
        dels = generate_deletions(ch.barriers, deletion_size) if deletions is None \
            else deletions.get(ch.name)
        if dels is None:
            continue
        cell_id = 0
        for ws, we, aws, awe in windows(ch.end, D):
            b0, b1 = ch.barrier_tree.equal_range(ws, we)
            if b0 == len(ch.barrier_tree):
                continue
            f1 = ch.feat_trees[0].equal_range(aws, awe)
            if f1[0] == len(ch.feat_trees[0]):
                continue
            f2 = ch.feat_trees[1].equal_range(aws, awe)
            if f2[0] == len(ch.feat_trees[1]):
                continue
            d0, d1 = dels.equal_range(ws, we)
            for k in range(d0, d1):
                twe, tawe = min(we, ch.end), min(awe, ch.end)
                ntarget = 0
                if density != 0.0:
                    npix1 = (twe - ws + bins - 1) // bins
                    npix2 = (D + bins - 1) // bins
                    ntarget = int(max(1.0, round(density * float(npix1 * npix2))))
                nlefs = int(round((twe - ws) / 1e6 * float(params.lefs_per_mbp)))
                tasks.append(Task(task_id, ch.name, ch.size, cell_id,
                                  target_simulation_epochs if ntarget == 0 else USIZE_MAX,
                                  ntarget, nlefs, int(dels.start[k]),
                                  int(dels.end[k] - dels.start[k]), ws, twe, aws, tawe,
                                  (b0, b1), f1, f2, ch.end))
                task_id += 1
                cell_id += 1
    return tasks


def window_barriers(ch: ChromInput, t: Task):
    """The task's barriers as simulate_window simulates them: the window's span of the
    chromosome's barriers, sorted by position (the ExtrusionBarriers constructor), with the
    deletion's barriers switched off (scheduler_perturbate.cpp:463-472, mgpu_apply_deletion).
    Returns (Barriers, number of barriers left active, counted over the whole span as the
    reference counts them).

    A barrier whose BED interval overlaps the window while its midpoint lies outside it (for
    records of a bp or two, a midpoint at the window end) stays in the reference's list but can
    block no LEF of the window; the device takes barriers inside the simulated range only, so
    such a barrier is left out of the task (its state draws no longer shift the cell's stream)."""
    b = S.sort_barriers(ch.barriers.subset(*t.barrier_range))
    inside = (b.pos >= t.window_start) & (b.pos < t.window_end)
    gone = (b.pos >= t.deletion_begin) & (b.pos < t.deletion_begin + t.deletion_size)
    total_active = len(b) - int(gone.sum())
    if not inside.all():
        b = S.Barriers(b.pos[inside], b.dir[inside], b.stp_active[inside], b.stp_inactive[inside])
    nact = C.c_uint64(0)
    L.check(L.lib().mgpu_apply_deletion(
        b.pos.ctypes.data_as(L.U64P), len(b), t.deletion_begin, t.deletion_size,
        b.stp_active.ctypes.data_as(L.DP), b.stp_inactive.ctypes.data_as(L.DP), C.byref(nact)))
    return b, total_active


def window_specs(chroms, tasks):
    """WindowSpec of each task for simulation.WindowJob, and each task's active barrier count."""
    by_name = {c.name: c for c in chroms}
    specs, active = [], []
    for t in tasks:
        ch = by_name[t.chrom]
        bars, nact = window_barriers(ch, t)
        target_epochs = t.num_target_epochs if t.num_target_contacts == 0 else USIZE_MAX
        specs.append(S.WindowSpec(t.chrom, t.chrom_size, t.window_start, t.window_end, bars,
                                  t.num_lefs, t.cell_id, t.num_target_contacts, target_epochs))
        active.append(nact)
    return specs, active


def task_line(t: Task) -> str:
    """A line of the task file (fmt::formatter<TaskPW>, simulation_impl.hpp:162-170)."""
    return "\t".join(str(x) for x in (
        t.id, t.chrom, t.cell_id, t.num_target_epochs, t.num_target_contacts, t.num_lefs,
        t.num_barriers, t.deletion_begin, t.deletion_size, t.window_start, t.window_end,
        t.active_window_start, t.active_window_end, t.feats1_range[1] - t.feats1_range[0],
        t.feats2_range[1] - t.feats2_range[0]))


def parse_task_line(line: str, chroms) -> Task:
    """TaskPW::from_string (simulation.cpp:640-700): the task's fields, its chromosome's
    barriers and features mapped to its windows, checked against the counts the line holds."""
    toks = line.rstrip("\n").split("\t")
    if len(toks) != 15:
        raise ValueError(f"task definition with {len(toks)} fields, expected 15: {line!r}")
    v = [int(x) if i != 1 else x for i, x in enumerate(toks)]
    by_name = {c.name: c for c in chroms}
    if v[1] not in by_name:
        raise ValueError(f'Unable to find a chromosome named "{v[1]}"')
    ch = by_name[v[1]]
    if len(ch.barriers) < v[6]:
        raise ValueError(f'Expected {v[6]} extrusion barriers for chromosome "{v[1]}". '
                         f"Found {len(ch.barriers)}")
    ws, we, aws, awe = v[9], v[10], v[11], v[12]
    b = ch.barrier_tree.equal_range(ws, we)
    f1 = ch.feat_trees[0].equal_range(aws, awe)
    f2 = ch.feat_trees[1].equal_range(aws, awe)
    if f1[1] - f1[0] != v[13] or f2[1] - f2[0] != v[14]:
        raise ValueError(f'Expected {v[13]} and {v[14]} features for chromosome "{v[1]}". '
                         f"Found {f1[1] - f1[0]} and {f2[1] - f2[0]}")
    return Task(v[0], v[1], ch.size, v[2], v[3], v[4], v[5], v[7], v[8], ws, we, aws, awe, b, f1,
                f2, ch.end)


def binomial_test(k: int, n: int) -> float:
    """stats::binomial_test<TWO_SIDED>(k, n, 0.5) (src/stats/tests_impl.hpp:16-85), which
    restates scipy.stats.binomtest's two-sided p-value."""
    from scipy.stats import binomtest
    return float(binomtest(int(k), int(n), 0.5, alternative="two-sided").pvalue)


def _fmt4g(x: float) -> str:
    return format(x, ".4G")


@dataclass
class PairQuery:
    task: int
    feat1: Feature
    feat2: Feature
    rel_bin1: int
    rel_bin2: int
    abs_bin1: int
    abs_bin2: int


def pair_queries(chroms, tasks, bin_size: int, diagonal_width: int):
    """The feature pairs simulate_window reads for each task (scheduler_perturbate.cpp:482-500):
    pairs (i, j) with j >= i over the two tracks, skipping pairs that both start at or past the
    partition point (active window start + D) except in the chromosome's last window; each
    feature at its centre's bin, relative to the window and absolute."""
    by_name = {c.name: c for c in chroms}
    out = []
    for ti, t in enumerate(tasks):
        ch = by_name[t.chrom]
        f1s = ch.feat_trees[0].data[t.feats1_range[0]:t.feats1_range[1]]
        f2s = ch.feat_trees[1].data[t.feats2_range[0]:t.feats2_range[1]]
        last_window = t.active_window_end == ch.end
        pp = t.active_window_start + diagonal_width
        for i, f1 in enumerate(f1s):
            c1 = (f1.start + f1.end + 1) // 2
            for j in range(i, len(f2s)):
                f2 = f2s[j]
                if not last_window and f1.start >= pp and f2.start >= pp:
                    continue
                c2 = (f2.start + f2.end + 1) // 2
                out.append(PairQuery(ti, f1, f2, (c1 - t.window_start) // bin_size,
                                     (c2 - t.window_start) // bin_size, c1 // bin_size,
                                     c2 // bin_size))
    return out


def bedpe_lines(tasks, queries, contacts, reference, active, bin_size: int):
    """The BEDPE lines of simulate_window (scheduler_perturbate.cpp:500-560) for the pair queries:
    contacts and reference are the block sums of the simulated and the reference band for each
    query; pairs with neither are left out. score = log2(contacts / reference) clamped to
    [-999, 999] (999 without reference contacts, -999 without simulated ones)."""
    lines = []
    for q, c, r in zip(queries, contacts.tolist(), reference.tolist()):
        if c == 0 and r == 0:
            continue
        t = tasks[q.task]
        if r == 0:
            score = 999.0
        elif c == 0:
            score = -999.0
        else:
            score = min(999.0, max(-999.0, math.log2(c / r)))
        sig = binomial_test(c, r + c)
        name = f"{q.feat1.name or 'none'};{q.feat2.name or 'none'}"
        lines.append("\t".join(str(x) for x in (
            q.feat1.chrom, q.abs_bin1 * bin_size, (q.abs_bin1 + 1) * bin_size,
            q.feat2.chrom, q.abs_bin2 * bin_size, (q.abs_bin2 + 1) * bin_size,
            name, _fmt4g(score), q.feat1.strand, q.feat2.strand, r, c, _fmt4g(sig),
            t.deletion_begin, t.deletion_begin + t.deletion_size, active[q.task],
            t.num_barriers, t.active_window_start, t.active_window_end, t.window_start,
            t.window_end, t.id)) + "\n")
    return lines


BEDPE_HEADER = ("#chrom1\tstart1\tend1\tchrom2\tstart2\tend2\tname\tscore\tstrand1\tstrand2\t"
                "reference_contacts\tcontacts\tsignificance\tdeletion_begin\tdeletion_end\t"
                "num_active_barriers\ttotal_num_barriers\tactive_window_start\t"
                "active_window_end\twindow_start\twindow_end\ttask_id\n")


def run(ctx: S.Context, params: L.MgpuParams, chroms, tasks, reference_bands, block_size: int = 9,
        mode: int = L.MODE_FAST):
    """Simulates every task in one job and returns (job, stats, BEDPE lines).
    reference_bands[t] is the reference matrix over task t's outer window (the band a
    ContactMatrixDense of the window holds, read_reference_contacts,
    scheduler_perturbate.cpp:81-91); its blocks are summed on the device as well."""
    specs, active = window_specs(chroms, tasks)
    job = S.WindowJob(ctx, params, specs, mode)
    st = job.run()
    qs = pair_queries(chroms, tasks, int(params.bin_size), int(params.diagonal_width))
    if not qs:
        return job, st, []
    ti = np.array([q.task for q in qs], np.uint64)
    r1 = np.array([q.rel_bin1 for q in qs], np.uint64)
    r2 = np.array([q.rel_bin2 for q in qs], np.uint64)
    contacts = job.block_sums(ti, r1, r2, block_size)
    reference = np.zeros(len(qs), np.uint32)
    for t in np.unique(ti).tolist():
        sel = np.nonzero(ti == t)[0]
        plan = job.plans[t]
        reference[sel] = S.band_block_sums(ctx, reference_bands[t], plan.nrows, plan.ncols,
                                           r1[sel], r2[sel], block_size)
    return job, st, bedpe_lines(tasks, qs, contacts, reference, active, int(params.bin_size))


def cool_name(prefix: str, t: Task) -> str:
    """The per-task file name replay writes (scheduler_perturbate.cpp:565-568)."""
    return (f"{prefix}_{t.id:06d}_{t.chrom}_window_{t.window_start}-{t.window_end}"
            f"_deletion_{t.deletion_begin}-{t.deletion_begin + t.deletion_size}.cool")


def replay(ctx: S.Context, params: L.MgpuParams, chroms, task_lines, prefix: str,
           task_filter=None, mode: int = L.MODE_FAST, genome=None):
    """Simulation::run_replay (scheduler_replay.cpp:39-154): the tasks of a task file (only the
    ids in task_filter when given) simulated again in one job, each written to its own .cool
    file with the window's matrix for its chromosome and empty matrices for the others
    (simulate_window with write_contacts_to_cooler, scheduler_perturbate.cpp:563-611). genome:
    [(name, size, start, end)] of every chromosome in genome order (default: the chroms given).
    Returns the paths written."""
    from .cooler import CoolerWriter
    tasks = []
    for line in task_lines:
        if not line.strip():
            continue
        t = parse_task_line(line, chroms)
        if task_filter is None or t.id in task_filter:
            tasks.append(t)
    if not tasks:
        return []
    specs, _ = window_specs(chroms, tasks)
    job = S.WindowJob(ctx, params, specs, mode)
    job.run()
    genome = genome or [(c.name, c.size, c.start, c.end) for c in chroms]
    longest = max(len(g[0]) for g in genome)
    paths = []
    for k, t in enumerate(tasks):
        band, _ = job.band(k)
        plan = job.plans[k]
        path = cool_name(prefix, t)
        with CoolerWriter(path, int(params.bin_size), longest) as w:
            for name, size, start, end in genome:
                if name == t.chrom:
                    w.write_chrom(name, t.window_start, t.window_end, size, band, plan.nrows,
                                  plan.ncols)
                else:
                    w.write_chrom(name, start, end, size)
        paths.append(path)
    job.close()
    return paths
