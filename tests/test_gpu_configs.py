"""Every benchmark configuration at its real size on the GPU.

  * genome (24 chromosomes of 40-250 Mbp, clustered barriers with Beta(8,2) occupancy, 512 cells
    each): fast mode against the oracle statistically on the smallest and the largest
    chromosome (128 of their cells, the 256-lane and the 1024-lane launch tiers), and every one
    of the 12 288 cells in one run with exact conservation checks;
  * stress (250 Mbp, 1 kb bins, 5 Mbp band, 10 000 LEFs, 12 000 barriers): fast mode against the
    oracle statistically on 32 cells (the hybrid LDS / device-memory placement), and all 2 048
    cells with exact conservation checks;
  * surrogate at its shipped shape (600 input bins, 8 x 128-channel encoder with dilations up to
    32, 100 x 400 head) with the trained weights on sweep inputs: forward and gradients against
    the CPU oracle;
  * the surrogate-accelerated sweep of 4 096 points with its 5 % judge.

Statistical tolerances come from the oracle itself: its cells are simulated as two halves, and
each statistic's split-half difference (at half the cells, so noisier than the comparison at
full size) sets the allowed fast-vs-oracle difference, with the floors of the small
configuration's test (test_gpu_fast.py) below which a difference is not looked at."""
import sys

import numpy as np
import pytest

import contact_stats as CS
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


def _fast_cells(ctx, wl, name, ncells):
    """Fast mode on cells [0, ncells) of one chromosome of the workload, with the per-cell
    targets of the full workload (so the cells are those of the full run)."""
    sim = S.Simulation(wl.options, ctx)
    names, chroms, bars, plans = sim.prepare(wl)
    i = names.index(name)
    p = plans[i]
    sub = S.ChromPlan(p.nlefs, p.nrows, p.ncols, p.targets[:ncells].copy(), p.target_epochs)
    job = S.Job(ctx, sim.params, [chroms[i]], [bars[i]], [sub], L.MODE_FAST)
    st = job.run()
    tun = job.tuning()
    band, missed = job.band(0)
    job.close()
    return band, missed, st, p, tun


def _oracle_halves(wl, name, ncells, nthreads=16):
    c = OC.transform(wl.options)
    rc = OC.to_rc(c)
    size = dict(wl.chroms)[name]
    bars = OC.prepare_barriers(wl.barriers_of(name), 0, size, c)
    p = S.plan_chrom(S.make_params(wl.options), S.chrom_view(name, size),
                     wl.options["num_cells"])
    ch = R.make_chrom(name, size, 0, size)
    h = ncells // 2
    t = np.ascontiguousarray(p.targets, np.uint64)
    a, ma, sa = R.simulate_cells(rc, ch, bars, p.nlefs, 0, h, t[:h], p.target_epochs, p.nrows,
                                 p.ncols, nthreads)
    b, mb, sb = R.simulate_cells(rc, ch, bars, p.nlefs, h, ncells, t[h:ncells], p.target_epochs,
                                 p.nrows, p.ncols, nthreads)
    return (a, ma, sa), (b, mb, sb), bars, p


def _compare(fband, fst, a, b, nrows, ncols, bars, bin_size, label):
    """Statistical comparison of a fast-mode band with the oracle's two halves (same cells,
    same targets). Returns a list of failures."""
    aband, bband = a[0], b[0]
    oband = aband.astype(np.uint32) + bband
    failures = []
    # contact decay P(s): log-spaced diagonal groups with >= 5000 oracle contacts (the rest pooled
    # into a tail group); allowed: max(10 %, 4 x the RMS split-half difference), tail 35 %
    _, osums, _ = CS.decay_groups(oband, nrows, ncols)
    keep = osums >= 5000
    po, osums, groups = CS.decay_groups(oband, nrows, ncols, keep)
    pf, _, _ = CS.decay_groups(fband, nrows, ncols, keep)
    pa, _, _ = CS.decay_groups(aband, nrows, ncols, keep)
    pb, _, _ = CS.decay_groups(bband, nrows, ncols, keep)
    full = osums >= 5000
    half = np.abs(pa - pb) / (0.5 * (pa + pb))
    rms_half = float(np.sqrt(np.mean(half[full] ** 2))) if full.any() else 0.0
    rel = np.abs(pf - po) / po
    tol = np.where(full, max(0.10, 4 * rms_half), max(0.35, 4 * rms_half))
    say(f"{label}: P(s) groups {[(int(x), int(y)) for x, y in groups]}")
    say(f"{label}: P(s) relative differences {np.round(rel, 3).tolist()} "
        f"(oracle split halves RMS {rms_half:.3f})")
    if (rel > tol).any():
        failures.append(f"P(s) {np.round(rel, 3).tolist()} over {np.round(tol, 3).tolist()}")
    # coarse-grained matrices (50 kb blocks): the fast-vs-oracle correlation (full samples) at
    # least the oracle's split-half correlation (half samples) less 0.01
    w = max(1, 50_000 // bin_size)
    r = CS.pearson(CS.block_sums(fband, nrows, ncols, w), CS.block_sums(oband, nrows, ncols, w))
    r_half = CS.pearson(CS.block_sums(aband, nrows, ncols, w), CS.block_sums(bband, nrows, ncols, w))
    say(f"{label}: block correlation {r:.4f} (oracle split halves {r_half:.4f})")
    if r < r_half - 0.01:
        failures.append(f"block correlation {r} vs split halves {r_half}")
    # barrier stripes and convergent-pair dots
    bins = (bars.pos // bin_size).astype(np.int64)
    for what, fn, floor in (("stripe", CS.barrier_stripe_enrichment, 0.10),
                            ("dot", CS.barrier_dot_enrichment, 0.30)):
        eo, ef = fn(oband, nrows, ncols, bins, bars.dir), fn(fband, nrows, ncols, bins, bars.dir)
        ea, eb = fn(aband, nrows, ncols, bins, bars.dir), fn(bband, nrows, ncols, bins, bars.dir)
        allowed = max(floor, 3 * abs(ea - eb) / eo)
        say(f"{label}: {what} enrichment oracle {eo:.3f} fast {ef:.3f} (halves {ea:.3f} / "
            f"{eb:.3f}, allowed {allowed:.3f})")
        if abs(ef - eo) / eo > allowed:
            failures.append(f"{what} {ef} vs {eo}")
    # epochs per cell: mean burn-in and total within max(5 %, 4 standard errors)
    ost = np.concatenate([a[2], b[2]])
    n = len(ost)
    for lab, col, fval in (("burn-in", 1, fst.burnin_epochs / fst.cells),
                           ("total", 0, fst.epochs / fst.cells)):
        x = ost[:, col].astype(np.float64)
        allowed = max(0.05 * x.mean(), 4.0 * x.std() * np.sqrt(2.0 / n))
        say(f"{label}: {lab} epochs per cell oracle {x.mean():.1f} fast {fval:.1f} "
            f"(allowed {allowed:.1f})")
        if abs(fval - x.mean()) > allowed:
            failures.append(f"{lab} epochs {fval} vs {x.mean()}")
    return failures


def _genome_chrom(which):
    wl = W.genome()
    by_len = sorted(wl.chroms, key=lambda c: c[1])
    return wl, (by_len[0] if which == "smallest" else by_len[-1])[0]


@pytest.mark.parametrize("which", ["smallest", "largest"])
def test_genome_chromosome_statistics_match_oracle(ctx, which):
    """128 cells of the genome workload's smallest (42 Mbp, 841 LEFs: four cells per CU) or
    largest chromosome (241 Mbp, 4 827 LEFs: one 1024-lane cell per CU), fast mode against the
    oracle, with the full workload's per-cell targets."""
    wl, name = _genome_chrom(which)
    ncells = 128
    say(f"genome {which} ({name}): oracle, {ncells} cells in two halves")
    a, b, bars, plan = _oracle_halves(wl, name, ncells)
    say(f"genome {which} ({name}): fast mode")
    fband, fmissed, fst, _, tun = _fast_cells(ctx, wl, name, ncells)
    say(f"genome {which}: launch tiers {list(tun.tier_threads[:tun.launches])}, "
        f"{fst.kernel_ms:.1f} ms")
    tot = int(plan.targets[:ncells].sum())
    assert fst.dropped == 0
    assert int(fband.sum()) + fmissed == tot == int(a[0].sum()) + a[1] + int(b[0].sum()) + b[1]
    c = OC.transform(wl.options)
    failures = _compare(fband, fst, a, b, plan.nrows, plan.ncols, bars, c["bin_size"],
                        f"genome {which}")
    assert not failures, failures


@pytest.mark.parametrize("which,ncells,kernel,slots", [("smallest", 16, "det_wave_kernel", 0),
                                                       ("largest", 8, "det_wave_slab_kernel", 0),
                                                       ("largest", 8, "det_wave_slab_kernel", 3)])
def test_genome_chromosome_deterministic_bit_exact(ctx, which, ncells, kernel, slots):
    """Deterministic mode at genome scale, against the oracle bit for bit (band, missed updates
    and epoch counts), with the full workload's per-cell targets: 16 cells of the genome
    workload's smallest chromosome (42 Mbp, 841 LEFs, clustered barriers with Beta(8,2)
    occupancy; one wavefront per cell, its 127 KB working set in LDS) and 8 cells of its largest
    (241 Mbp, 4 827 LEFs; the cell's arrays, about 0.5 MB, in a device-memory slot, its
    engine-word ring in LDS), also with the slab capped at 3 slots, so the 8 cells run as three
    launches over the same slots (mgpu_debug_det_slab_slots)."""
    wl, name = _genome_chrom(which)
    c = OC.transform(wl.options)
    rc = OC.to_rc(c)
    size = dict(wl.chroms)[name]
    bars = OC.prepare_barriers(wl.barriers_of(name), 0, size, c)
    sim = S.Simulation(wl.options, ctx)
    names, chroms, dbars, plans = sim.prepare(wl)
    i = names.index(name)
    p = plans[i]
    t = np.ascontiguousarray(p.targets[:ncells], np.uint64)
    say(f"genome deterministic ({name}, {p.nlefs} LEFs): oracle, {ncells} cells")
    ch = R.make_chrom(name, size, 0, size)
    oband, omissed, ost = R.simulate_cells(rc, ch, bars, p.nlefs, 0, ncells, t, p.target_epochs,
                                           p.nrows, p.ncols, 16)
    say(f"genome deterministic ({name}): device")
    sub = S.ChromPlan(p.nlefs, p.nrows, p.ncols, t.copy(), p.target_epochs)
    ctx.check(L.lib().mgpu_debug_det_slab_slots(ctx.ptr, slots))
    try:
        job = S.Job(ctx, sim.params, [chroms[i]], [dbars[i]], [sub], L.MODE_DETERMINISTIC)
    finally:
        ctx.check(L.lib().mgpu_debug_det_slab_slots(ctx.ptr, 0))
    assert job.kernel() == kernel
    if slots:
        assert job.tuning().grid == slots
    st = job.run()
    band, missed = job.band(0)
    cs = job.cell_stats()
    job.close()
    say(f"genome deterministic ({name}): {st.kernel_ms:.1f} ms, {st.epochs} epochs")
    assert int(st.epochs) == int(ost[:, 0].sum())
    # per cell: epochs, burn-in epochs, contacts and engine words consumed
    assert np.array_equal(cs, ost), (cs, ost)
    assert missed == omissed
    assert np.array_equal(band, oband), int(np.count_nonzero(band != oband))


def test_slab_cells_optional_outputs_bit_exact(ctx):
    """1D LEF occupancy and the state log from cells whose arrays sit in device memory: 4 cells
    of the genome workload's largest chromosome (4 827 LEFs) with both outputs on, against the
    oracle's run with the same outputs, bit for bit (band, occupancy, every record)."""
    wl, name = _genome_chrom("largest")
    ncells = 4
    opts = dict(wl.options, track_1d_lef_position=1, log_model_internal_state=1)
    c = OC.transform(opts)
    rc = OC.to_rc(c)
    size = dict(wl.chroms)[name]
    bars = OC.prepare_barriers(wl.barriers_of(name), 0, size, c)
    sim = S.Simulation(opts, ctx)
    names, chroms, dbars, plans = sim.prepare(wl)
    i = names.index(name)
    p = plans[i]
    t = np.ascontiguousarray(p.targets[:ncells], np.uint64)
    say(f"slab outputs ({name}): oracle, {ncells} cells")
    ch = R.make_chrom(name, size, 0, size)
    occ = np.zeros((size + c["bin_size"] - 1) // c["bin_size"], np.uint64)
    oband, omissed, ost, orecs = R.simulate_cells(rc, ch, bars, p.nlefs, 0, ncells, t,
                                                  p.target_epochs, p.nrows, p.ncols, 4,
                                                  occupancy=occ, state_log_cap=20_000)
    say("slab outputs: device")
    sub = S.ChromPlan(p.nlefs, p.nrows, p.ncols, t.copy(), p.target_epochs)
    job = S.Job(ctx, sim.params, [chroms[i]], [dbars[i]], [sub], L.MODE_DETERMINISTIC)
    assert job.kernel() == "det_wave_slab_kernel"
    job.set_state_log(20_000)
    job.run()
    band, missed = job.band(0)
    docc = job.occupancy(0)
    drecs, nrec = job.state_log()
    job.close()
    assert missed == omissed and np.array_equal(band, oband)
    assert np.array_equal(docc[:len(occ)], occ), int(np.count_nonzero(docc[:len(occ)] != occ))
    assert drecs is not None and nrec == len(orecs) == int(ost[:, 0].sum())
    for f in ("epoch", "burnin", "num_active_lefs", "active_barriers", "stalls_rev", "stalls_fwd",
              "stalls_both", "lef_bar", "primary", "secondary", "avg_loop_size"):
        assert np.array_equal(drecs[f], orecs[f]), f


def _decay_is_decreasing(band, nrows, ncols, slack=1.05):
    p, _ = CS.log_binned_decay(band, nrows, ncols)
    return bool(np.all(p[1:] <= p[:-1] * slack)), p


def test_genome_workload_full_size_conserves_contacts(ctx):
    """All 24 x 512 cells of the genome workload in one run (the benchmark's step): every
    chromosome's band plus its out-of-band count equals its planned contacts, no contact is
    dropped, the padding element stays zero, and contact frequency falls with distance."""
    wl = W.genome()
    sim = S.Simulation(wl.options, ctx)
    names, chroms, bars, plans = sim.prepare(wl)
    job = S.Job(ctx, sim.params, chroms, bars, plans, L.MODE_FAST)
    st = job.run()
    say(f"genome full size: {st.cells} cells, {st.kernel_ms:.1f} ms, launches {job.launches()}")
    assert st.cells == 24 * 512 and st.dropped == 0
    total = 0
    for i, name in enumerate(names):
        band, missed = job.band(i)
        p = plans[i]
        want = int(p.targets.sum())
        total += want
        assert band[p.nrows * p.ncols] == 0, name
        assert int(band.sum(dtype=np.uint64)) + missed == want, name
        ok, dec = _decay_is_decreasing(band, p.nrows, p.ncols)
        assert ok, (name, dec)
    job.close()
    assert st.contacts == total


def test_stress_statistics_match_oracle(ctx):
    """32 cells of the stress workload (250 Mbp, 1 kb bins, 5 000 diagonals, 10 000 LEFs, 12 000
    barriers: working sets too large for LDS, run hybrid), fast mode against the oracle, with
    the full workload's per-cell targets (610 352 contacts per cell)."""
    wl = W.stress()
    ncells = 32
    say(f"stress: oracle, {ncells} cells in two halves")
    a, b, bars, plan = _oracle_halves(wl, "chrL", ncells)
    say("stress: fast mode")
    fband, fmissed, fst, _, tun = _fast_cells(ctx, wl, "chrL", ncells)
    say(f"stress: launch tiers {list(tun.tier_threads[:tun.launches])}, {fst.kernel_ms:.1f} ms")
    tot = int(plan.targets[:ncells].sum())
    assert fst.dropped == 0
    assert int(fband.sum(dtype=np.uint64)) + fmissed == tot
    assert int(a[0].sum(dtype=np.uint64)) + a[1] + int(b[0].sum(dtype=np.uint64)) + b[1] == tot
    failures = _compare(fband, fst, a, b, plan.nrows, plan.ncols, bars, 1000, "stress")
    assert not failures, failures


def test_stress_cell_deterministic_bit_exact(ctx):
    """Deterministic mode on the stress workload's configuration: one cell of the 250 Mbp
    chromosome (1 kb bins, 5 000 diagonals, 10 000 LEFs, 12 000 barriers; 4 160 epochs) with the
    full workload's target, its arrays (1 MB) in a device-memory slot, against the oracle bit for
    bit: band, missed updates, epochs, burn-in epochs, contacts and the 194 M engine words
    consumed. About 22 s on the device."""
    wl = W.stress()
    name, size = wl.chroms[0]
    c = OC.transform(wl.options)
    rc = OC.to_rc(c)
    bars = OC.prepare_barriers(wl.barriers_of(name), 0, size, c)
    sim = S.Simulation(wl.options, ctx)
    names, chroms, dbars, plans = sim.prepare(wl)
    p = plans[0]
    t = np.ascontiguousarray(p.targets[:1], np.uint64)
    say("stress deterministic: oracle, 1 cell")
    ch = R.make_chrom(name, size, 0, size)
    oband, omissed, ost = R.simulate_cells(rc, ch, bars, p.nlefs, 0, 1, t, p.target_epochs,
                                           p.nrows, p.ncols, 1)
    say("stress deterministic: device")
    sub = S.ChromPlan(p.nlefs, p.nrows, p.ncols, t.copy(), p.target_epochs)
    job = S.Job(ctx, sim.params, chroms, dbars, [sub], L.MODE_DETERMINISTIC)
    assert job.kernel() == "det_wave_slab_kernel"
    st = job.run()
    band, missed = job.band(0)
    cs = job.cell_stats()
    job.close()
    say(f"stress deterministic: {st.kernel_ms:.0f} ms, {st.epochs} epochs")
    assert np.array_equal(cs, ost), (cs, ost)
    assert missed == omissed
    assert np.array_equal(band, oband), int(np.count_nonzero(band != oband))


def test_stress_workload_full_size_conserves_contacts(ctx):
    """All 2 048 cells of the stress workload: the 5 GB band plus its out-of-band count equals
    the planned 1.25 billion contacts, none dropped, and contact frequency falls with distance."""
    wl = W.stress()
    out, st = S.Simulation(wl.options, ctx).run_simulate(wl, L.MODE_FAST)
    band, missed, plan = out["chrL"]
    say(f"stress full size: {st.cells} cells, {st.kernel_ms:.1f} ms, {st.contacts} contacts")
    assert st.cells == 2048 and st.dropped == 0
    assert band[plan.nrows * plan.ncols] == 0
    assert int(band.sum(dtype=np.uint64)) + missed == int(plan.targets.sum()) == st.contacts
    ok, dec = _decay_is_decreasing(band, plan.nrows, plan.ncols)
    assert ok, dec


def test_surrogate_at_shipped_shape_matches_oracle(ctx):
    """The shipped network (default shape, trained weights) on two windows of sweep inputs:
    forward within 3 % of the output range (correlation > 0.999) and every parameter tensor's
    gradient within cosine 0.999 and 2 % relative norm of the CPU oracle's (fp64, run on the
    same bf16-rounded weights and inputs), the first encoder layer's included."""
    import os

    from modlefork_amd import surrogate as SG
    from modlefork_amd import sweep as Sw
    from oracle import surrogate_ref as SR

    net = SG.Surrogate(ctx, SG.default_config(max_batch=2))
    SG.load_params(net, os.path.join(os.path.dirname(SG.__file__), "data",
                                     "surrogate_weights.npy"))
    c = SR.SgConfig()
    p = net.params()
    layout = Sw.fixed_layout()
    x = Sw.point_inputs(layout, Sw.sample_points(1, seed=42)[0])[5:7].astype(np.float32)
    pr = SR.unflatten(c, SG.bf16_round(p).astype(np.float64))
    xr = SG.bf16_round(x).astype(np.float64)
    dev = net.forward(x)
    ref, _ = SR.forward(c, pr, xr)
    err = np.abs(dev - ref).max() / np.abs(ref).max()
    corr = np.corrcoef(dev.ravel(), ref.ravel())[0, 1]
    say(f"shipped surrogate forward: max error {err:.4f} of the output range, corr {corr:.6f}")
    assert err < 0.03 and corr > 0.999
    # targets 30 % below the prediction (plus noise): a residual with a systematic part, as in
    # training. With targets equal to the prediction plus zero-mean noise, every gradient is a
    # sum of cancelling terms, and bf16 activations alone move it by 25 % (measured on the
    # oracle with bf16 rounding at each stored tensor); with this residual, by below 1 %.
    rng = np.random.default_rng(8)
    y = (0.7 * ref + 0.05 * rng.standard_normal(ref.shape)).astype(np.float32)
    loss, g = net.gradients(x, y)
    rl, rg, _ = SR.loss_and_grads(c, pr, xr, y.astype(np.float64))
    say(f"shipped surrogate loss device {loss:.6f} oracle {rl:.6f}")
    assert abs(loss - rl) < 0.02 * rl
    k = 0
    for name, sh in SR.param_shapes(c):
        n = int(np.prod(sh))
        ga, gb = g[k:k + n].astype(np.float64), rg[k:k + n]
        k += n
        nb = np.linalg.norm(gb)
        cos = float(ga @ gb / (np.linalg.norm(ga) * nb + 1e-30))
        rel = float(np.linalg.norm(ga - gb) / (nb + 1e-30))
        say(f"  {name:10s} cos {cos:.5f} rel {rel:.4f}")
        assert cos > 0.999 and rel < 0.02, name  # measured: rel <= 0.005 in every tensor
    net.close()


def test_full_sweep_judges_and_accepts_the_surrogate(ctx):
    """The 4 096-point sweep with the shipped weights: 205 points (5 %) judged by simulation, the
    surrogate accepted (at most a fifth of the judged points fail), every failing point replaced
    by its simulation, the mean judged correlation at least 0.95 and the sweep faster than
    simulating every point."""
    import os

    from modlefork_amd import surrogate as SG
    from modlefork_amd import sweep as Sw

    layout = Sw.fixed_layout()
    # eight whole chromosomes per forward batch, as the benchmark runs it
    net = SG.Surrogate(ctx, SG.default_config(
        max_batch=8 * len(Sw.window_starts(layout.size // 5000))))
    SG.load_params(net, os.path.join(os.path.dirname(SG.__file__), "data",
                                     "surrogate_weights.npy"))
    pts = Sw.sample_points(4096)
    res, bands = Sw.run_sweep(ctx, net, layout, pts, log=say, rate_sample=32)
    net.close()
    say(f"sweep: {res.sweep_rate:.1f} matrices/s, plain simulation {res.sim_rate:.2f}")
    assert len(res.judged) == 205 and not res.fell_back
    assert len(res.replaced) <= 0.2 * len(res.judged)
    assert sorted(bands) == sorted(res.replaced)
    assert np.nanmean(res.correlations) >= 0.95
    assert all(not c >= 0.95 for k, c in zip(res.judged, res.correlations) if k in bands)
    assert res.sweep_rate > res.sim_rate
    say(f"raw 5 kb correlation over all 99 off-diagonals: {np.nanmean(res.raw_5kb):.4f}")


def test_training_windows_as_configured(ctx):
    """The surrogate's training data as its configuration states it: isolated 2 Mbp windows
    (simulated with their 500 kb of context on each side), 0-60 uniform barriers, occupancy
    U(0.3, 0.95), LEF density U(5, 40)/Mbp, processivity U(100, 600) kb, 256 cells per window,
    fast mode. 512 of the 20 000 windows here (bench.py --training-data 20000 runs the whole set,
    98.5 s on one MI355X): input tracks carry each window's barriers, every window registers
    contacts and contact frequency falls with distance."""
    from modlefork_amd import windows as WN

    specs = WN.sample_specs(512, seed=3)
    x, y, kernel_s = WN.simulate_windows(ctx, specs, ncells=256, contact_density=10.0, chunk=2048)
    say(f"training windows: 512 x 256 cells, kernel {kernel_s:.2f} s")
    assert x.shape == (512, WN.IN_BINS, WN.IN_CH) and y.shape == (512, WN.NDIAG, WN.OUT_BINS)
    for k in (0, 100, 511):
        assert int(x[k, :, 0].sum()) == len(specs[k].records)
    assert np.all(y.sum(axis=(1, 2)) > 0)
    per_diag = y.mean(axis=(0, 2))
    assert per_diag[1] > per_diag[20] > per_diag[90] > 0


def test_attenuation_correction_is_calibrated(ctx):
    """The judge's correction for the simulation's own noise, checked with a prediction whose
    agreement with the expected band is known: a simulation of the same point with 16 times the
    judge's cells and contacts (other seeds). Spearman-Brown gives its reliability
    q16 = 16 q / (1 + 15 q) from the judge's reliability q, so on each judged diagonal its raw
    correlation with the judge's simulation should be sqrt(q * q16) and the corrected one
    sqrt(q16), close to 1. Over three sweep points: the mean corrected correlation (unclipped)
    lies within 0.03 of the expected mean of sqrt(q16), the raw one within 0.02 of the mean of
    sqrt(q * q16); a prediction of the wrong point (another density and processivity) fails the
    0.95 bar."""
    from modlefork_amd import sweep as Sw

    layout = Sw.fixed_layout()
    pts = [Sw.Point(12.0, 200e3, 0.9), Sw.Point(25.0, 400e3, 0.7), Sw.Point(35.0, 550e3, 1.0)]
    wrong = Sw.Point(6.0, 120e3, 0.6)
    for k, p in enumerate(pts):
        halves, _ = Sw.simulate_points(ctx, layout, [p, p], 256, 15.0, seed=3,
                                       stored_diagonals=100, name_ids=[10 * k, 10 * k + 1])
        big, _ = Sw.simulate_points(ctx, layout, [p, wrong], 8192, 480.0, seed=4,
                                    stored_diagonals=100, name_ids=[10 * k + 2, 10 * k + 3])
        a, b = halves
        full = a + b
        ds = Sw.well_sampled(full, Sw.MIN_COUNT)
        h = Sw.diag_correlations(a, b, diagonals=ds)
        q = 2 * h / (1 + h)
        keep = q >= Sw.MIN_RELIABILITY
        ds = [D for D, kk in zip(ds, keep) if kk]
        q = q[keep]
        r = Sw.diag_correlations(big[0], full, diagonals=ds)
        q16 = 16 * q / (1 + 15 * q)
        corrected = float(np.mean(r / np.sqrt(q)))
        say(f"point {k}: {len(ds)} judged diagonals, reliability {q.mean():.3f}, raw "
            f"{r.mean():.4f} (expected {np.mean(np.sqrt(q * q16)):.4f}), corrected {corrected:.4f} "
            f"(expected {np.mean(np.sqrt(q16)):.4f})")
        assert len(ds) >= 5
        assert abs(corrected - np.mean(np.sqrt(q16))) < 0.03
        assert abs(r.mean() - np.mean(np.sqrt(q * q16))) < 0.02
        c_wrong, _, _ = Sw.judge(big[1], a, b)
        say(f"point {k}: the wrong point's simulation as prediction scores {c_wrong:.4f}")
        assert not c_wrong >= 0.95
