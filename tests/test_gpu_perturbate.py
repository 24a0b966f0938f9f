"""perturbate / replay windows on the GPU (mgpu_job_create_windows, mgpu_job_block_sums,
mgpu_band_block_sums): deterministic-mode window tasks bit for bit against the oracle (window
bounds, deleted barriers, per-task cell ids), the block readout against the oracle and the
reference's known answers, fast-mode windows conserving contacts with deletions that act, the
BEDPE readout, and replay reproducing perturbate's bands into .cool files."""
import os
import sys

import numpy as np
import pytest

import perturb_case as PC
import unit_kat_replay as K
from modlefork_amd import _lib as L
from modlefork_amd import perturbate as P
from modlefork_amd import simulation as S
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


def _oracle_window(ch, t, c_rc):
    """The oracle's band of one task: the window's barriers (overlapping BED intervals, in the
    chromosome's tree order) sorted by position, the deletion's barriers set to (0, 1), one
    cell with the task's id over [window_start, window_end)."""
    imp = ch.barriers
    hit = np.nonzero((imp.bed_start < t.window_end) & (imp.bed_end > t.window_start))[0]
    rows = [(int(imp.bed_start[i]), int(imp.bed_end[i]), int(imp.pos[i]), int(imp.dir[i]),
             float(imp.stp_active[i]), float(imp.stp_inactive[i]))
            for i in range(hit[0], hit[-1] + 1)]
    b = OC.sort_barriers(rows)
    inside = (b.pos >= t.window_start) & (b.pos < t.window_end)  # see window_barriers
    pos, dirs = b.pos[inside], b.dir[inside]
    sa, si = b.stp_active[inside].copy(), b.stp_inactive[inside].copy()
    gone = (pos >= t.deletion_begin) & (pos < t.deletion_begin + t.deletion_size)
    sa[gone], si[gone] = 0.0, 1.0
    bars = R.Barriers(pos, dirs, sa, si)
    rch = R.make_chrom(t.chrom, t.chrom_size, t.window_start, t.window_end)
    ncols = -(-(t.window_end - t.window_start) // PC.OPTIONS["bin_size"])
    nrows = min(PC.OPTIONS["diagonal_width"] // PC.OPTIONS["bin_size"], ncols)
    band, missed, st = R.simulate_cells(c_rc, rch, bars, t.num_lefs, t.cell_id, t.cell_id + 1,
                                        np.array([t.num_target_contacts], np.uint64),
                                        t.num_target_epochs, nrows, ncols)
    return band, missed, st, int(gone.sum())


def test_deterministic_windows_match_the_oracle(ctx):
    """Six tasks of the plan (first and middle windows, several deletions) in one deterministic
    job: every task's band, missed count and epochs equal the oracle's."""
    params, chroms, _ = PC.build()
    tasks = P.plan(params, chroms)
    pick = tasks[:3] + tasks[len(tasks) // 2:len(tasks) // 2 + 3]
    say(f"deterministic windows: {len(pick)} of {len(tasks)} tasks")
    specs, active = P.window_specs(chroms, pick)
    job = S.WindowJob(ctx, params, specs, L.MODE_DETERMINISTIC)
    st = job.run()
    c_rc = OC.to_rc(OC.transform(PC.OPTIONS))
    epochs = 0
    for k, t in enumerate(pick):
        band, missed = job.band(k)
        ob, om, ost, ngone = _oracle_window(chroms[0], t, c_rc)
        say(f"task {t.id}: window [{t.window_start}, {t.window_end}) deletion "
            f"[{t.deletion_begin}, {t.deletion_begin + t.deletion_size}) ({ngone} barriers), "
            f"cell {t.cell_id}, {int(band.sum())} contacts")
        assert ngone >= 1 and active[k] == t.num_barriers - ngone
        assert np.array_equal(band, ob) and missed == om, t.id
        epochs += int(ost[:, 0].sum())
    assert st.epochs == epochs and st.dropped == 0
    job.close()


def test_block_sums_on_the_device(ctx):
    """mgpu_band_block_sums on the reference's "CMatrix get w/ block" matrix gives its three
    answers; mgpu_job_block_sums on simulated window bands equals the oracle's block sums for
    random pixels, edges included, at block sizes 1, 5 and 9."""
    case = K.load()["cmatrix"]["CMatrix get w/ block"]
    band = K.block_band(case)
    q = np.array(case["block_checks"], np.uint64)
    got = S.band_block_sums(ctx, band, case["nrows"], case["ncols"], q[:, 0], q[:, 1], 5)
    assert got.tolist() == q[:, 3].tolist()

    params, chroms, _ = PC.build()
    tasks = P.plan(params, chroms)[:4]
    specs, _ = P.window_specs(chroms, tasks)
    job = S.WindowJob(ctx, params, specs, L.MODE_FAST)
    job.run()
    rng = np.random.default_rng(4)
    for bs in (1, 5, 9):
        ti = rng.integers(0, len(tasks), 2000).astype(np.uint64)
        rows, cols, want = [], [], []
        bands = [job.band(k) for k in range(len(tasks))]
        for t in ti.tolist():
            plan = job.plans[t]
            r = int(rng.integers(0, plan.ncols))
            c = int(np.clip(r + rng.integers(-plan.nrows - 10, plan.nrows + 10), 0, plan.ncols - 1))
            if rng.random() < 0.1:
                r, c = 0, int(rng.integers(0, 3))
            rows.append(r)
            cols.append(c)
            want.append(R.band_get_block(bands[t][0], plan.nrows, plan.ncols, r, c, bs))
        got = job.block_sums(ti, rows, cols, bs)
        assert got.tolist() == want, bs
    with pytest.raises(L.MgpuError):
        job.block_sums([0], [0], [0], 4)
    job.close()


def test_fast_windows_conserve_contacts_and_deletions_act(ctx):
    """Every task of the plan in one fast-mode job: band + missed + dropped equals each task's
    target. A window whose barriers are all deleted loses the contact enrichment at its
    barriers that the same window keeps without deletions (both simulated here)."""
    params, chroms, _ = PC.build()
    tasks = P.plan(params, chroms)
    specs, _ = P.window_specs(chroms, tasks)
    job = S.WindowJob(ctx, params, specs, L.MODE_FAST)
    st = job.run()
    say(f"fast windows: {len(tasks)} tasks, {st.kernel_ms:.1f} ms, {st.contacts} contacts")
    assert st.dropped == 0
    for k, t in enumerate(tasks):
        band, missed = job.band(k)
        assert int(band.sum()) + missed == t.num_target_contacts, t.id
    job.close()

    # one window, 64 cells each: intact barriers against all barriers deleted
    t = tasks[len(tasks) // 2]
    intact = S.sort_barriers(chroms[0].barriers.subset(*t.barrier_range))
    gone = S.Barriers(intact.pos, intact.dir, np.zeros(len(intact)), np.ones(len(intact)))
    specs = []
    for bars in (intact, gone):
        for cell in range(64):
            specs.append(S.WindowSpec(t.chrom, t.chrom_size, t.window_start, t.window_end, bars,
                                      t.num_lefs, cell, t.num_target_contacts, P.USIZE_MAX))
    job = S.WindowJob(ctx, params, specs, L.MODE_FAST)
    job.run()
    plan = job.plans[0]
    sums = [sum(job.band(k)[0].astype(np.int64) for k in range(h * 64, h * 64 + 64))
            for h in range(2)]
    job.close()
    bins = ((intact.pos - t.window_start) // PC.OPTIONS["bin_size"]).astype(np.int64)
    import contact_stats as CS
    e_in = CS.barrier_stripe_enrichment(sums[0], plan.nrows, plan.ncols, bins, intact.dir)
    e_gone = CS.barrier_stripe_enrichment(sums[1], plan.nrows, plan.ncols, bins, intact.dir)
    say(f"stripe enrichment at the window's barriers: intact {e_in:.3f}, deleted {e_gone:.3f}")
    assert e_in > 1.05 and abs(e_gone - 1.0) < 0.5 * (e_in - 1.0)


def test_perturbate_readout_and_replay(ctx, tmp_path):
    """run() on the plan's first window with the undeleted simulation of each task's window as
    the reference matrix: BEDPE lines with 22 fields whose score and significance follow from
    their contact columns; then replay of three task lines in deterministic mode writes .cool
    files whose bands equal perturbate's bands for those tasks (same window, deletion and cell),
    with empty matrices for the other chromosome."""
    params, chroms, _ = PC.build()
    tasks = P.plan(params, chroms)
    first = [t for t in tasks if t.window_start == tasks[0].window_start
             and t.active_window_start == tasks[0].active_window_start]
    # reference matrices: each task's window without deletions, 1 cell, same cell id
    ref_specs = []
    for t in first:
        bars = S.sort_barriers(chroms[0].barriers.subset(*t.barrier_range))
        ref_specs.append(S.WindowSpec(t.chrom, t.chrom_size, t.window_start, t.window_end, bars,
                                      t.num_lefs, t.cell_id, t.num_target_contacts, P.USIZE_MAX))
    rj = S.WindowJob(ctx, params, ref_specs, L.MODE_FAST)
    rj.run()
    refs = [rj.band(k)[0] for k in range(len(first))]
    rj.close()
    job, st, lines = P.run(ctx, params, chroms, first, refs, block_size=9)
    say(f"perturbate readout: {len(first)} tasks, {len(lines)} BEDPE lines")
    assert lines
    for ln in lines:
        f = ln.rstrip("\n").split("\t")
        assert len(f) == 22
        ref, con = int(f[10]), int(f[11])
        want = 999.0 if ref == 0 else (-999.0 if con == 0 else np.log2(con / ref))
        assert abs(float(f[7]) - want) <= 5e-4 * max(1.0, abs(want))
        assert abs(float(f[12]) - P.binomial_test(con, con + ref)) <= 5e-4 * P.binomial_test(con, con + ref) + 1e-300
    job.close()

    chosen = tasks[:2] + tasks[-1:]
    specs, _ = P.window_specs(chroms, chosen)
    pj = S.WindowJob(ctx, params, specs, L.MODE_DETERMINISTIC)
    pj.run()
    want = [pj.band(k)[0] for k in range(len(chosen))]
    pj.close()
    prefix = str(tmp_path / "rpl")
    paths = P.replay(ctx, params, chroms, [P.task_line(t) + "\n" for t in tasks],
                     prefix, task_filter={t.id for t in chosen}, mode=L.MODE_DETERMINISTIC)
    assert [os.path.basename(p) for p in paths] == [os.path.basename(P.cool_name(prefix, t))
                                                    for t in chosen]
    from modlefork_amd import cooler as CL
    from test_cool_writer import H5, H5READ, _rebuild

    if not os.path.exists(CL.LIB_PATH) or not os.path.exists(H5READ):
        pytest.skip("HDF5 is not available: the .cool writer is not built")
    bs = PC.OPTIONS["bin_size"]
    for p, t, w in zip(paths, chosen, want):
        h = H5(p)
        assert h.strings("chroms/name") == ["chrP", "chrQ"]
        ncols = -(-(t.window_end - t.window_start) // bs)
        nrows = min(PC.OPTIONS["diagonal_width"] // bs, ncols)
        p1, p2, cnt = h.i64("pixels/bin1_id"), h.i64("pixels/bin2_id"), h.i32("pixels/count")
        assert np.all(p2 < 6_000_000 // bs)  # chrQ holds no pixels
        got = _rebuild(p1 - t.window_start // bs, p2 - t.window_start // bs, cnt, 0, nrows, ncols)
        assert np.array_equal(got[:-1], w[:-1]), p
