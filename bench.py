"""Benchmark of the simulate hot path on one MI355X.

A step is one pass of the hot path over one batch: every (chromosome, cell) task of the workload
simulated to its contact target, with the inputs and the bands resident in HBM. Prints one JSON
line (cells/s as `value`, contacts/s and epochs/s beside it), with the roofline of the simulation
kernel and the CPU oracle timed on this host's cores on a bounded sample of the same workload.
A "surrogate" object beside them reports the surrogate-accelerated parameter sweep: matrices/s of
the network and of the plain simulation, the judged correlation and the network's MFMA rate.

Fast mode runs one persistent launch per size tier at the tiers' default workgroup sizes (1024,
512 and 256 lanes per cell). --autotune first times every size on a sample of each tier's cells
(an eighth of them, and at least as many as the chip keeps resident) and keeps the fastest per
tier: on the genome workload it keeps 1024 and 512 lanes for the two large tiers and finds 256
and 512 lanes within noise for the small one, so the default run is the tuned one; the rocprof
counters under profiles/ are keyed by the launch shape they were taken on. The timed steps run
as replays of a HIP graph of the run (mgpu_job_capture: the band clears and every launch,
captured once after the warm-up). Graph replays carry no per-launch events, so one eager run
after the timed steps gives each launch's duration for the roofline (`eager_run` reports its
rate beside `value`).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload genome|small|stress]
                    [--mode fast|det] [--cpu-cells K] [--cpu-threads T] [--no-stress]
                    [--sweep-points P | --no-sweep]
                    [--autotune]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# L2 atomics: 16 channels per XCD x 8 XCDs, one 32-bit atomic per channel per 2.4 GHz clock
ATOMIC_PEAK_PER_S = 128 * 2.4e9


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(name, ncells):
    from modlefork_amd import workloads as W

    if name == "genome":
        return W.genome(ncells=ncells or 512)
    if name == "small":
        return W.small(ncells=ncells or 64)
    if name == "stress":
        return W.stress(ncells=ncells or 2048)
    raise SystemExit(f"unknown workload {name}")


def launch_bytes(ls) -> int:
    """Algorithmic bytes of one simulation launch: 8 B per registered contact (a read and a
    write of one int32 pixel), 21 B per barrier-table entry each cell reads (position, direction
    and transition probabilities) and 32 B of engine state per cell. The fast mode's working sets
    stay on chip (LDS) in the measured configurations, so no LEF state spills."""
    return 8 * int(ls.contacts) + 21 * int(ls.barriers) + 32 * int(ls.cells)


def spill_bytes(cell_stats, plans, bars) -> int:
    """The spill term of the algorithmic bytes for cells whose carried state lives in device
    memory: per cell-epoch a read and a write of 20 B per LEF (positions, binding epoch, ranks)
    and 1 B per barrier (its state). cell_stats: [cells, 4] in job order (epochs first)."""
    total, k = 0, 0
    for p, b in zip(plans, bars):
        n = len(p.targets)
        ep = int(cell_stats[k:k + n, 0].sum())
        total += ep * 2 * (20 * int(p.nlefs) + len(b.pos))
        k += n
    return total


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo, as lscpu reports it)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, options, cells_per_chrom, nthreads, single_cells=1):
    """The oracle on this host's cores, on a stratified sample of the workload: cells
    0 .. cells_per_chrom - 1 of every chromosome (the same fraction of each, so the sample has the
    workload's mix of chromosome sizes), run as one task per cell from a shared queue on
    `nthreads` worker threads, the way simulate_worker takes tasks (scheduler_simulate.cpp:
    186-284); each chromosome's cells add into one shared band. The workload's wall time is
    extrapolated from the summed per-cell times divided by the threads (a full run balances
    thousands of cells over the threads, so its tail is negligible; the sample's own wall time,
    which includes its tail, is reported beside it). A 1-thread run of `single_cells` cells per
    chromosome gives the per-core rate."""
    import threading
    from concurrent.futures import ThreadPoolExecutor

    from oracle import config as OC
    from oracle import refcpu as R

    R.build()
    c = OC.transform(options)
    rc = OC.to_rc(c)
    cfg = R.make_config(rc)
    ncells = options["num_cells"]
    per_chrom = []
    total_cells = 0
    for name, size in wl.chroms:
        bars = OC.prepare_barriers(wl.barriers_of(name), 0, size, c)
        if len(bars.pos) == 0:
            continue
        total_cells += ncells
        nl = R.lib().rc_num_lefs(c["number_of_lefs_per_mbp"], size)
        ncols = R.lib().rc_band_ncols(c["bin_size"], size)
        nrows = R.lib().rc_band_nrows(c["diagonal_width"], c["bin_size"], size)
        t = np.zeros(ncells, np.uint64)
        te = R.lib().rc_plan_targets(C.byref(cfg), ncells, nl, nrows * ncols,
                                     t.ctypes.data_as(C.POINTER(C.c_uint64)))
        per_chrom.append(dict(ch=R.make_chrom(name, size, 0, size), bars=bars, nl=nl, nrows=nrows,
                              ncols=ncols, t=t, te=te, band=None, lock=threading.Lock()))

    def run_cell(k, cell):
        pc = per_chrom[k]
        with pc["lock"]:
            if pc["band"] is None:
                pc["band"] = np.zeros(pc["nrows"] * pc["ncols"] + 1, np.uint32)
        t0 = time.perf_counter()
        R.simulate_cells(rc, pc["ch"], pc["bars"], pc["nl"], cell, cell + 1,
                         pc["t"][cell:cell + 1], pc["te"], pc["nrows"], pc["ncols"], 1,
                         band=pc["band"])
        return time.perf_counter() - t0

    # per-core rate: single_cells cells of every chromosome on this thread alone
    single = [run_cell(k, cell) for k in range(len(per_chrom)) for cell in range(single_cells)]
    tasks = [(k, cell) for k in range(len(per_chrom))
             for cell in range(single_cells, single_cells + cells_per_chrom)]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=nthreads) as ex:
        times = list(ex.map(lambda kc: run_cell(*kc), tasks))
    wall = time.perf_counter() - t0
    for pc in per_chrom:
        pc["band"] = None
    rate = len(tasks) * nthreads / sum(times)
    return {
        "value": rate,
        "unit": "cells/s",
        "cores": nthreads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "per_core_cells_per_s": len(single) / sum(single),
        "sample_wall_cells_per_s": len(tasks) / wall,
        "sample": f"oracle (C restatement of the reference) on {len(tasks)} of the workload's "
                  f"{total_cells} cells ({cells_per_chrom} of each chromosome), one task per "
                  f"cell on {nthreads} threads; rate = cells x threads / summed per-cell time "
                  f"({sum(times):.1f} thread-s; the sample's own wall {wall:.1f} s gives "
                  f"{len(tasks) / wall:.1f} cells/s); per core from {len(single)} cells on one "
                  f"thread ({sum(single):.1f} s)",
    }


def stress_leg(ctx, mode, nthreads, no_cpu):
    """The large-state stress workload (one 250 Mbp chromosome, 1 kb bins, 5 Mbp band, 10 000
    LEFs, 12 000 barriers, 2048 cells; working sets run hybrid, LDS and device memory): one timed
    run after a warm-up, and the oracle on 2 cells per thread of it."""
    from modlefork_amd import simulation as S
    from modlefork_amd import workloads as W

    wl = W.stress()
    sim = S.Simulation(wl.options, ctx)
    names, chroms, bars, plans = sim.prepare(wl)
    job = S.Job(ctx, sim.params, chroms, bars, plans, mode)
    job.run()
    t0 = time.perf_counter()
    st = job.run()
    wall = time.perf_counter() - t0
    job.close()
    out = {"workload": wl.name, "cells": int(st.cells), "cells_per_sec": st.cells / wall,
           "contacts_per_sec": st.contacts / wall, "epochs_per_sec": st.epochs / wall,
           "kernel_ms": float(st.kernel_ms), "dropped": int(st.dropped)}
    log(f"stress: {st.cells} cells in {wall:.2f} s ({st.cells / wall:.1f} cells/s)")
    if not no_cpu:
        cpu = cpu_baseline(wl, wl.options, 2 * nthreads, nthreads)
        out["cpu_baseline"] = cpu
        out["speedup_vs_cpu"] = out["cells_per_sec"] / cpu["value"]
    return out


def deterministic_leg(ctx, nthreads, no_cpu, genome_cpu=None, phases=False):
    """Deterministic mode (the reference's draws, bands bit-exact against the oracle) on the
    small workload: 20 Mbp, 400 LEFs, 200 barriers, 64 cells, one wavefront per cell. Three
    timed runs after a warm-up, the same layout with 512 cells (the chip keeps about that many
    wave-parallel cells resident, so this is the mode's throughput), and the oracle on the 64
    cells on this host's threads. Then the genome workload's first 64 cells of each of its 24
    chromosomes with their full-workload targets (1 536 cells, 841-4 827 LEFs: the cells' arrays
    in device-memory slots, engine-word rings in LDS), one timed run after a warm-up, against
    the CPU rate the genome workload's own baseline measured (genome_cpu). With `phases`, one
    more genome-sample run with the per-phase timers gives the summed wavefront time, and so how
    many wavefronts were busy on average over the kernel. Last, the whole genome workload
    (12 288 cells) once, with its contacts checked per chromosome."""
    from modlefork_amd import _lib as L
    from modlefork_amd import simulation as S
    from modlefork_amd import workloads as W

    out = {}
    for ncells in (64, 512):
        wl = W.small(ncells=ncells)
        sim = S.Simulation(wl.options, ctx)
        names, chroms, bars, plans = sim.prepare(wl)
        job = S.Job(ctx, sim.params, chroms, bars, plans, L.MODE_DETERMINISTIC)
        kernel = job.kernel()
        job.run()
        runs = 3 if ncells == 64 else 1
        t0 = time.perf_counter()
        for _ in range(runs):
            st = job.run()
        wall = (time.perf_counter() - t0) / runs
        words = int(job.cell_stats()[:, 3].sum())
        job.close()
        rec = {"cells": int(st.cells), "cells_per_sec": st.cells / wall,
               "contacts_per_sec": st.contacts / wall, "epochs_per_sec": st.epochs / wall,
               "draws_per_sec": words / wall, "draws_per_cell_epoch": words / st.epochs,
               "kernel_ms": float(st.kernel_ms), "kernel": kernel}
        log(f"deterministic {wl.name}: {st.cells} cells in {wall * 1e3:.1f} ms "
            f"({st.cells / wall:.1f} cells/s, {kernel})")
        if ncells == 64:
            out.update({"workload": wl.name, **rec})
            if not no_cpu:
                cpu = cpu_baseline(wl, wl.options, 64 - 4, nthreads, single_cells=4)
                out["cpu_baseline"] = cpu
                out["speedup_vs_cpu"] = out["cells_per_sec"] / cpu["value"]
        else:
            out["batch_512_cells"] = rec

    wl = W.genome()
    sim = S.Simulation(wl.options, ctx)
    names, chroms, bars, plans = sim.prepare(wl)
    sub = [S.ChromPlan(p.nlefs, p.nrows, p.ncols, np.ascontiguousarray(p.targets[:64]),
                       p.target_epochs) for p in plans]
    job = S.Job(ctx, sim.params, chroms, bars, sub, L.MODE_DETERMINISTIC)
    kernel = job.kernel()
    job.run()
    t0 = time.perf_counter()
    st = job.run()
    wall = time.perf_counter() - t0
    words = int(job.cell_stats()[:, 3].sum())
    busy = None
    if phases:
        ctx.phase_times(True)
        pst = job.run()
        ph = ctx.phase_times(False, deterministic=True).astype(np.float64)
        wave_s = float(ph.sum()) / 1e8
        busy = {"wave_seconds": wave_s, "kernel_ms": float(pst.kernel_ms),
                "mean_busy_wavefronts": wave_s / (float(pst.kernel_ms) / 1e3),
                "us_per_cell_epoch": wave_s / float(pst.epochs) * 1e6,
                "phase_shares": {L.DET_PHASE_NAMES[k]: float(v / ph.sum())
                                 for k, v in enumerate(ph) if v}}
        log(f"deterministic genome sample phases: {wave_s:.1f} s of wavefront time in "
            f"{pst.kernel_ms:.1f} ms ({busy['mean_busy_wavefronts']:.0f} wavefronts busy on "
            f"average, {busy['us_per_cell_epoch']:.0f} us per cell-epoch)")
        for k, v in enumerate(ph):
            if v:
                log(f"  {L.DET_PHASE_NAMES[k]:28s} {100.0 * v / ph.sum():5.1f} %")
    job.close()
    if phases:
        # the chain that bounds the sample: the largest chromosome's 64 cells on their own
        big = max(range(len(plans)), key=lambda i: sub[i].nlefs)
        jb = S.Job(ctx, sim.params, [chroms[big]], [bars[big]], [sub[big]], L.MODE_DETERMINISTIC)
        jb.run()
        sb = jb.run()
        jb.close()
        busy["largest_chromosome_alone"] = {"chromosome": names[big], "lefs": int(sub[big].nlefs),
                                            "cells": int(sb.cells),
                                            "kernel_ms": float(sb.kernel_ms),
                                            "epochs_per_cell": float(sb.epochs) / float(sb.cells)}
        log(f"deterministic genome sample: {names[big]} ({sub[big].nlefs} LEFs) alone, "
            f"{sb.cells} cells: {sb.kernel_ms:.1f} ms")
    rec = {"workload": "genome-24chroms-5kb, first 64 cells per chromosome",
           "cells": int(st.cells), "cells_per_sec": st.cells / wall,
           "contacts_per_sec": st.contacts / wall, "epochs_per_sec": st.epochs / wall,
           "draws_per_sec": words / wall, "draws_per_cell_epoch": words / st.epochs,
           "kernel_ms": float(st.kernel_ms), "kernel": kernel}
    log(f"deterministic genome sample: {st.cells} cells in {wall * 1e3:.1f} ms "
        f"({st.cells / wall:.1f} cells/s, {kernel})")
    if genome_cpu:
        rec["cpu_cells_per_sec"] = genome_cpu["value"]
        rec["speedup_vs_cpu"] = rec["cells_per_sec"] / genome_cpu["value"]
    if busy:
        rec["phases"] = busy
    out["genome_sample"] = rec

    # the whole genome workload (24 x 512 cells), one run (the kernel has just run above), with
    # every chromosome's contacts checked: band sum + missed == its cells' targets
    job = S.Job(ctx, sim.params, chroms, bars, plans, L.MODE_DETERMINISTIC)
    t0 = time.perf_counter()
    st = job.run()
    wall = time.perf_counter() - t0
    spill = spill_bytes(job.cell_stats(), plans, bars)
    ls = job.launch_stats()
    alg = sum(launch_bytes(x) for x in ls) + spill
    conserved = True
    for i in range(len(plans)):
        band, missed = job.band(i)
        conserved &= int(band.sum(dtype=np.uint64)) + missed == int(plans[i].targets.sum())
    full = {"workload": "genome-24chroms-5kb-512cells", "cells": int(st.cells),
            "cells_per_sec": st.cells / wall, "contacts_per_sec": st.contacts / wall,
            "epochs_per_sec": st.epochs / wall, "kernel_ms": float(st.kernel_ms),
            "kernel": job.kernel(), "launches": int(job.tuning().launches),
            "contacts_conserved": bool(conserved), "dropped": int(st.dropped),
            "roofline": {"bound": "hbm", "algorithmic_bytes": alg, "spill_bytes": spill,
                         "achieved": alg / (st.kernel_ms / 1000.0) / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s",
                         "frac": alg / (st.kernel_ms / 1000.0) / 1e9 / HBM_PEAK_GBS}}
    job.close()
    log(f"deterministic genome workload: {st.cells} cells in {wall * 1e3:.1f} ms "
        f"({st.cells / wall:.1f} cells/s, contacts conserved: {conserved})")
    if genome_cpu:
        full["cpu_cells_per_sec"] = genome_cpu["value"]
        full["speedup_vs_cpu"] = full["cells_per_sec"] / genome_cpu["value"]
    out["genome_full"] = full
    return out


def training_data_leg(ctx, nwin: int, density: float = 10.0):
    """The surrogate's training data as the training configuration states it: nwin isolated
    windows of 2 Mbp (simulated with their 500 kb of context on each side, 3 Mbp, 5 kb bins),
    0-60 uniform barriers, occupancy U(0.3, 0.95), LEF density U(5, 40)/Mbp, processivity
    U(100, 600) kb, 256 cells per window, fast mode, contact density `density`; chunks of 2 048
    windows per job, bands read back and cut to the 100-diagonal window targets."""
    from modlefork_amd import windows as WN

    specs = WN.sample_specs(nwin, seed=3)
    t0 = time.perf_counter()
    x, y, kernel_s = WN.simulate_windows(ctx, specs, ncells=256, contact_density=density,
                                         chunk=2048)
    wall = time.perf_counter() - t0
    cells = nwin * 256
    log(f"training data: {nwin} windows x 256 cells in {wall:.1f} s ({nwin / wall:.1f} "
        f"windows/s, kernel {kernel_s:.1f} s)")
    return {"windows": nwin, "cells_per_window": 256, "contact_density": density,
            "wall_s": wall, "kernel_s": kernel_s, "windows_per_sec": nwin / wall,
            "cells_per_sec": cells / wall, "inputs": list(x.shape), "targets": list(y.shape)}


def perturbate_inputs(wl, params):
    """ChromInput list of a perturbate workload (barriers in interval-tree order, features)."""
    from modlefork_amd import perturbate as P
    from modlefork_amd import simulation as S

    chroms = []
    for name, size in wl.chroms:
        imp = S.import_barriers(params, wl.barriers_of(name), 0, size)
        tracks = wl.features.get(name, [[], []])
        f = [[P.Feature(name, a, b, n, 0.0, st) for a, b, n, st in tr] for tr in tracks]
        chroms.append(P.ChromInput(name, size, 0, size, imp, f[0], f[1]))
    return chroms


def perturbate_cpu_baseline(wl, tasks, chroms, ntasks, nthreads):
    """The oracle on a stratified sample of the perturbate tasks (every k-th), one task per
    thread from a shared queue: per-task times give tasks x threads / summed time."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import config as OC
    from oracle import refcpu as R

    R.build()
    c = OC.transform(wl.options)
    rc = OC.to_rc(c)
    by_name = {ch.name: ch for ch in chroms}
    pick = tasks[::max(1, len(tasks) // ntasks)][:ntasks]

    def run_task(t):
        imp = by_name[t.chrom].barriers
        i0, i1 = t.barrier_range
        rows = [(int(imp.bed_start[i]), int(imp.bed_end[i]), int(imp.pos[i]), int(imp.dir[i]),
                 float(imp.stp_active[i]), float(imp.stp_inactive[i])) for i in range(i0, i1)]
        b = OC.sort_barriers(rows)
        keep = (b.pos >= t.window_start) & (b.pos < t.window_end)
        sa, si = b.stp_active[keep].copy(), b.stp_inactive[keep].copy()
        gone = (b.pos[keep] >= t.deletion_begin) & (b.pos[keep] < t.deletion_begin + t.deletion_size)
        sa[gone], si[gone] = 0.0, 1.0
        bars = R.Barriers(b.pos[keep], b.dir[keep], sa, si)
        ncols = -(-(t.window_end - t.window_start) // c["bin_size"])
        nrows = min(-(-c["diagonal_width"] // c["bin_size"]), ncols)
        t0 = time.perf_counter()
        R.simulate_cells(rc, R.make_chrom(t.chrom, t.chrom_size, t.window_start, t.window_end),
                         bars, t.num_lefs, t.cell_id, t.cell_id + 1,
                         np.array([t.num_target_contacts], np.uint64), t.num_target_epochs,
                         nrows, ncols, 1)
        return time.perf_counter() - t0

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=nthreads) as ex:
        times = list(ex.map(run_task, pick))
    wall = time.perf_counter() - t0
    return {"value": len(pick) * nthreads / sum(times), "unit": "tasks/s", "cores": nthreads,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": f"oracle on {len(pick)} of the {len(tasks)} tasks (every "
                      f"{max(1, len(tasks) // ntasks)}-th), one task per thread on {nthreads} "
                      f"threads; rate = tasks x threads / summed per-task time "
                      f"({sum(times):.1f} thread-s; the sample's wall {wall:.1f} s)"}


def perturbate_leg(ctx, nthreads, no_cpu):
    """modle perturbate's simulations and block readout on the perturb workload (one 30 Mbp
    chromosome, 3 Mbp band, one 10 kb deletion per barrier, windows of 12 Mbp): every
    (window, deletion) task in one fast-mode job, timed after a warm-up run, then the 9 x 9 block
    sums of every feature pair of every task on the device."""
    from modlefork_amd import _lib as L
    from modlefork_amd import perturbate as P
    from modlefork_amd import simulation as S
    from modlefork_amd import workloads as W

    wl = W.perturb()
    params = S.make_params(wl.options)
    chroms = perturbate_inputs(wl, params)
    tasks = P.plan(params, chroms)
    specs, _ = P.window_specs(chroms, tasks)
    job = S.WindowJob(ctx, params, specs, L.MODE_FAST)
    job.run()
    t0 = time.perf_counter()
    st = job.run()
    wall = time.perf_counter() - t0
    qs = P.pair_queries(chroms, tasks, int(params.bin_size), int(params.diagonal_width))
    ti = np.array([q.task for q in qs], np.uint64)
    r1 = np.array([q.rel_bin1 for q in qs], np.uint64)
    r2 = np.array([q.rel_bin2 for q in qs], np.uint64)
    t1 = time.perf_counter()
    job.block_sums(ti, r1, r2, 9)
    tq = time.perf_counter() - t1
    job.close()
    out = {"workload": wl.name, "tasks": len(tasks), "tasks_per_sec": len(tasks) / wall,
           "contacts_per_sec": st.contacts / wall, "epochs_per_sec": st.epochs / wall,
           "kernel_ms": float(st.kernel_ms), "dropped": int(st.dropped),
           "pair_queries": len(qs), "block_sums_ms": 1000.0 * tq}
    log(f"perturbate: {len(tasks)} tasks in {wall:.2f} s ({len(tasks) / wall:.1f} tasks/s), "
        f"{len(qs)} block sums in {1000 * tq:.1f} ms")
    if not no_cpu:
        cpu = perturbate_cpu_baseline(wl, tasks, chroms, 2 * nthreads, nthreads)
        out["cpu_baseline"] = cpu
        out["speedup_vs_cpu"] = out["tasks_per_sec"] / cpu["value"]
    return out


def surrogate_flops(cfg) -> float:
    """Algorithmic FLOPs of one surrogate forward pass over one window (multiply-adds of the real
    channels counted as 2 FLOPs; padding and the pairwise sum not counted)."""
    L, C, K = cfg.in_len, cfg.enc_ch, cfg.enc_k
    f = 2.0 * L * C * cfg.in_ch * K + (cfg.enc_layers - 1) * 2.0 * L * C * C * K
    f += 2 * 2.0 * cfg.out_len * C * cfg.head_ch
    pix = cfg.ndiag * cfg.out_len
    f += (cfg.head_layers - 1) * 2.0 * pix * cfg.head_ch * cfg.head_ch * 9 + 2.0 * pix * cfg.head_ch * 9
    return f


def surrogate_leg(ctx, npoints: int):
    """The surrogate-accelerated sweep (4096 points over LEF density, processivity and an
    occupancy scale on a fixed 20 Mbp layout, 5 % judged by 512-cell simulations), with the
    surrogate's MFMA throughput."""
    from modlefork_amd import surrogate as SG
    from modlefork_amd import sweep as Sw

    weights = os.path.join(ROOT, "modlefork_amd", "data", "surrogate_weights.npy")
    if not os.path.exists(weights):
        log("no trained surrogate weights: sweep leg skipped")
        return None
    layout = Sw.fixed_layout()
    nwin = len(Sw.window_starts(layout.size // 5000))
    batch = 8 * nwin  # eight whole chromosomes per forward batch
    net = SG.Surrogate(ctx, SG.default_config(max_batch=batch))
    SG.load_params(net, weights)
    res, _ = Sw.run_sweep(ctx, net, layout, Sw.sample_points(npoints), log=log)
    ms = net.predict_timed(batch, 10) / 10
    flops = surrogate_flops(net.cfg)
    tflops = flops * batch / (ms / 1000.0) / 1e12
    dev_rate = batch / (ms / 1000.0) / nwin
    net.close()
    corr = np.array(res.correlations, dtype=np.float64)
    return {
        "metric": "surrogate vs simulation matrices/s at stated correlation",
        "points": res.npoints, "judged": len(res.judged), "cells_per_judged_point": 512,
        "corr_threshold": 0.95,
        "corr_mean": float(np.nanmean(corr)), "corr_min": float(np.nanmin(corr)),
        "raw_corr_mean": float(np.nanmean(res.raw_correlations)),
        "raw_corr_5kb_all_diagonals_mean": float(np.nanmean(res.raw_5kb)),
        "judge_contact_density": 30.0,  # also the plain simulation's, whose rate is the baseline
        "judge_reliability": float(np.nanmean(res.reliabilities)),
        "passed": int(np.sum(corr >= 0.95)), "replaced": len(res.replaced),
        "surrogate_rejected": res.fell_back,
        "sweep_wall_s": res.sweep_s, "surrogate_wall_s": res.surrogate_s,
        "judge_wall_s": res.judge_s,
        "sim_matrices_per_s": res.sim_rate, "surrogate_matrices_per_s": res.surrogate_rate,
        "surrogate_device_matrices_per_s": dev_rate,
        "sweep_matrices_per_s": res.sweep_rate,
        "speedup_inference": res.surrogate_rate / res.sim_rate,
        "speedup_sweep": res.sweep_rate / res.sim_rate,
        "mfma": {"bound": "mfma", "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
                 "frac": tflops / 2500.0, "flops_per_window": flops, "dtype": "bf16"},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="genome")
    ap.add_argument("--mode", default=os.environ.get("MGPU_BENCH_MODE", "fast"))
    ap.add_argument("--ncells", type=int, default=0)
    ap.add_argument("--cpu-cells", type=int, default=0,
                    help="cells per chromosome of the CPU baseline's sample (0: 4 per thread, or "
                         "2 per thread on the stress workload)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0: this process's CPUs, at most 16: the host's "
                         "share of one GPU)")
    ap.add_argument("--no-perturbate", action="store_true",
                    help="skip the perturbate leg (windows with deletions and block readout)")
    ap.add_argument("--no-stress", action="store_true",
                    help="skip the stress leg (250 Mbp, 1 kb bins, 2048 cells: one GPU run and its "
                         "CPU baseline)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-deterministic", action="store_true",
                    help="skip the deterministic-mode leg (small workload, bit-exact mode)")
    ap.add_argument("--sweep-points", type=int, default=4096)
    ap.add_argument("--no-sweep", action="store_true", help="skip the surrogate sweep leg")
    ap.add_argument("--threads", type=int, default=0,
                    help="fast mode: workgroup size per cell (256, 512 or 1024; 0 = default)")
    ap.add_argument("--autotune", action="store_true",
                    help="time every workgroup size per size tier first and keep the fastest")
    ap.add_argument("--autotune-cells", type=int, default=0,
                    help="cells the autotuner times each workgroup size on (0: an eighth)")
    ap.add_argument("--phases", action="store_true",
                    help="after the timed steps, run one more step with the epoch-phase timers on "
                         "and print each phase's share of the time to stderr (fast mode, or the "
                         "wave-parallel deterministic kernel)")
    ap.add_argument("--training-data", type=int, default=0, metavar="WINDOWS",
                    help="also simulate the surrogate's training windows as its configuration "
                         "states them (isolated 2 Mbp windows, 256 cells each; 20000 for the "
                         "whole set)")
    ap.add_argument("--write-cool", default="", metavar="PATH",
                    help="also write the bands of one run to PATH as .cool and time it (outside "
                         "the timed steps; the file is removed afterwards)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (reference for HBM counters)")
    args = ap.parse_args()

    if args.copy_calibration:
        # a plain copy kernel of known size: reads and writes MIB MiB once each, so its
        # FETCH_SIZE / WRITE_SIZE counters give the scale of the HBM byte counters
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "launches": 3}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the simulate path shards cells across GPUs with no "
              "exchange until a final band reduction, and this build runs on a single device.",
              file=sys.stderr)
        sys.exit(2)

    from modlefork_amd import _lib as L
    from modlefork_amd import simulation as S

    wl = build_workload(args.workload, args.ncells)
    mode = L.MODE_FAST if args.mode == "fast" else L.MODE_DETERMINISTIC
    sim = S.Simulation(wl.options, S.Context(0))
    # genome import: barrier records to sorted device-ready tables, the task split (the
    # reference's import_genome and run_simulate planning), reported beside the timed steps
    tp = time.perf_counter()
    names, chroms, bars, plans = sim.prepare(wl)
    prepare_ms = 1000.0 * (time.perf_counter() - tp)
    log(f"workload {wl.name}: {len(chroms)} chromosomes, "
        f"{sum(len(p.targets) for p in plans)} cells, mode {args.mode} "
        f"(import and planning {prepare_ms:.1f} ms)")
    job = S.Job(sim.ctx, sim.params, chroms, bars, plans, mode)
    autotuned = None
    if args.threads:
        job.set_tuning(args.threads)
    elif mode == L.MODE_FAST and args.autotune:
        # the autotuner times every workgroup size on a sample of each size tier's cells and
        # keeps the fastest per tier (untimed: before the warmup)
        total = sum(len(p.targets) for p in plans)
        sample = args.autotune_cells or max(1, total // 8)
        sizes, tms = job.autotune(sample)
        autotuned = {"sample_cells": sample, "threads": sizes,
                     "ms_by_threads": {str(k): round(float(v), 1) for k, v in tms.items()}}
        log(f"autotune on {sample} cells: {sizes} (ms by size {autotuned['ms_by_threads']})")
    tuning = job.tuning()
    launches = [{"threads": int(tuning.tier_threads[k]), "grid": int(tuning.tier_grid[k])}
                for k in range(int(tuning.launches))]
    shape = "+".join(str(x["threads"]) for x in launches)
    log(f"launches (threads per cell, grid): {[(x['threads'], x['grid']) for x in launches]}")

    for i in range(args.warmup):
        st = job.run()
        log(f"warmup {i}: kernel {st.kernel_ms:.1f} ms")
    # The timed steps replay a HIP graph of the run (its band clears and launches, captured
    # once after the warm-up); every replay is a whole run of the workload.
    job.capture()
    job.run()  # the graph's first replay, untimed
    kernel_ms = []
    t0 = time.perf_counter()
    stats = None
    for i in range(args.steps):
        stats = job.run()  # synchronizes the stream before and after
        kernel_ms.append(stats.kernel_ms)
        log(f"step {i}: graph replay, {stats.kernel_ms:.1f} ms on the device, "
            f"{stats.epochs} epochs")
    wall = time.perf_counter() - t0
    job.release_graph()
    # one eager run after them: graph replays carry no per-launch events, so the launches'
    # own durations (the roofline below) come from this run
    te = time.perf_counter()
    est = job.run()
    wall_e = time.perf_counter() - te
    ls = job.launch_stats()
    eager = {"ms": 1000.0 * wall_e, "kernel_ms": float(est.kernel_ms),
             "cells_per_sec": est.cells / wall_e}
    log(f"eager run: {eager['ms']:.1f} ms ({eager['cells_per_sec']:.1f} cells/s)")
    if args.phases:
        det = mode == L.MODE_DETERMINISTIC
        sim.ctx.phase_times(True)
        job.run()
        ph = sim.ctx.phase_times(False, deterministic=det).astype(np.float64)
        pnames = L.DET_PHASE_NAMES if det else L.PHASE_NAMES
        log(f"epoch phases (share of cell time, "
            f"{'one wavefront' if det else f'{shape} threads'} per cell):")
        for k, v in enumerate(ph):
            if v:
                log(f"  {pnames[k]:28s} {100.0 * v / ph.sum():5.1f} %  ({v / 1e8:.3f} s summed)")
    ms_per_step = 1000.0 * wall / args.steps
    cells = int(stats.cells)
    contacts = int(stats.contacts)
    epochs = int(stats.epochs)
    value = cells * args.steps / wall

    # the rate a caller sees when the bands are handed over in host memory (as
    # mgpu_simulate_chrom does): one more run, then every band copied to pageable host memory
    t1 = time.perf_counter()
    job.run()
    t2 = time.perf_counter()
    band_bytes = 0
    results = {}
    for c in range(len(chroms)):
        band, missed = job.band(c)
        band_bytes += band.nbytes
        if args.write_cool:
            results[names[c]] = (band, missed, plans[c])
    t3 = time.perf_counter()
    host_bands = {"cells_per_s": cells / (t3 - t1), "readback_ms": 1000.0 * (t3 - t2),
                  "band_bytes": band_bytes,
                  "readback_gb_per_s": band_bytes / (t3 - t2) / 1e9}
    log(f"with the bands read back to host memory: {host_bands['cells_per_s']:.1f} cells/s "
        f"({host_bands['readback_ms']:.1f} ms for {band_bytes / 1e9:.2f} GB)")
    cool = None
    if args.write_cool:
        # the output step (Simulation::write_contacts_to_disk) on the bands just read back
        from modlefork_amd import cooler as CL

        t4 = time.perf_counter()
        CL.write_cool(args.write_cool, wl.chroms, results, int(sim.params.bin_size),
                      metadata={"workload": wl.name})
        cool_ms = 1000.0 * (time.perf_counter() - t4)
        cool = {"ms": cool_ms, "file_bytes": os.path.getsize(args.write_cool),
                "band_bytes": band_bytes}
        os.remove(args.write_cool)
        log(f".cool write: {cool_ms:.1f} ms, {cool['file_bytes'] / 1e9:.2f} GB file")
        del results

    # roofline of the simulation kernel. Fast mode launches it once per working-set size tier,
    # each launch on a stream of its own, so the launches of one run overlap on the chip (a
    # tier's workgroups take the CUs the tier before it leaves); the kernel's figure is therefore
    # all launches' algorithmic bytes over the run's device time (the span from the first
    # launch's start to the last one's end, HIP events on the job's stream), with each launch's
    # own bytes and duration beside it
    all_bytes = sum(launch_bytes(x) for x in ls)
    if job.kernel() == "det_wave_slab_kernel":  # carried state in device memory
        all_bytes += spill_bytes(job.cell_stats(), plans, bars)
    span_ms = float(est.kernel_ms)  # the eager run: launches timed by their own events
    achieved = all_bytes / (span_ms / 1000.0) / 1e9
    per_launch = [{"threads": int(x.threads), "grid": int(x.grid), "cells": int(x.cells),
                   "contacts": int(x.contacts), "ms": float(x.ms),
                   "algorithmic_bytes": launch_bytes(x)} for x in ls]

    # measured HBM bytes, LDS and atomic figures of each launch from the committed rocprof
    # counter runs (taken for one launch shape; other shapes report none)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{args.mode}_{args.workload}.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            prof = json.load(fh)
        hb = prof.get("hbm_bytes_per_launch", {})
        if prof.get("launch_shape") == shape and all(str(x.threads) in hb for x in ls):
            traffic = float(sum(hb[str(x.threads)] for x in ls))
    lds = atomics = None
    cfile = os.path.join(ROOT, "profiles", f"counters_{args.mode}_{args.workload}.json")
    if os.path.exists(cfile):
        with open(cfile) as fh:
            cprof = json.load(fh)
        pl = cprof.get("per_launch", {})
        if cprof.get("launch_shape") == shape and all(str(x.threads) in pl for x in ls):
            lds = {"per_launch": {}, "peak": "LDS busy on every cycle of all 256 CUs "
                                              "(SQ_LDS_IDX_ACTIVE)"}
            count = 0.0
            for x in ls:
                dv = pl[str(x.threads)]["derived"]
                lds["per_launch"][str(x.threads)] = {
                    "util": dv["lds_util"], "bank_conflict_ratio": dv["lds_bank_conflict_ratio"],
                    "instructions": dv["lds_instructions"],
                    "valu_issue_util": dv["valu_issue_util"],
                    "wait_fraction": dv["wait_fraction"]}
                count += dv["atomics"]
            per_s = count / (span_ms / 1000.0)
            atomics = {"per_s": per_s, "count": count, "peak_per_s": ATOMIC_PEAK_PER_S,
                       "frac": per_s / ATOMIC_PEAK_PER_S}

    cpu = None
    nthreads = args.cpu_threads or min(len(os.sched_getaffinity(0)), 16)
    if not args.no_cpu:
        k = args.cpu_cells or (2 if args.workload == "stress" else 4) * nthreads
        log(f"cpu baseline: {k} cells per chromosome on {nthreads} threads")
        cpu = cpu_baseline(wl, wl.options, k, nthreads)

    line = {
        "metric": "cells/sec + contacts/sec at 1 GPU",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64/f64" if mode == L.MODE_DETERMINISTIC else "u32/f64",
        "data": "synthetic",
        "config": {"workload": wl.name, "chromosomes": len(chroms), "cells": cells,
                   "mode": "deterministic" if mode == L.MODE_DETERMINISTIC else "fast",
                   "bin_size": int(sim.params.bin_size),
                   "diagonal_width": int(sim.params.diagonal_width),
                   "lefs_per_mbp": float(sim.params.lefs_per_mbp), "parallelism": "none",
                   "launches": launches, "autotuned": autotuned},
        "contacts_per_sec": contacts * args.steps / wall,
        "with_host_readback": host_bands,
        "eager_run": eager,
        "genome_import_ms": prepare_ms,
        "cool_write": cool,
        "epochs_per_sec": epochs * args.steps / wall,
        "kernel_ms": float(np.mean(kernel_ms)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": job.kernel(),
                     "ms": span_ms, "algorithmic_bytes": all_bytes,
                     "launches": per_launch},
        "lds": lds,
        "atomics": atomics,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu"] = value / cpu["value"]
    job.close()
    if not args.no_stress and args.workload == "genome":
        line["stress"] = stress_leg(sim.ctx, mode, nthreads, args.no_cpu)
    if not args.no_perturbate and args.workload == "genome":
        line["perturbate"] = perturbate_leg(sim.ctx, nthreads, args.no_cpu)
    if not args.no_deterministic and args.workload == "genome":
        line["deterministic"] = deterministic_leg(sim.ctx, nthreads, args.no_cpu, cpu,
                                                  phases=args.phases)
    if args.training_data:
        line["training_data"] = training_data_leg(sim.ctx, args.training_data)
    if not args.no_sweep:
        log(f"surrogate sweep: {args.sweep_points} points")
        line["surrogate"] = surrogate_leg(sim.ctx, args.sweep_points)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
