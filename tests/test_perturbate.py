"""Host side of perturbate / replay (modlefork_amd.perturbate) and the block readout, on the CPU:
the reference's unsafe_get_block known answers on the oracle, barrier import order against the
oracle, deletions, windows, the task plan and the task file. The device runs are in
test_gpu_perturbate.py."""
import numpy as np
import pytest

import perturb_case as PC
import unit_kat_replay as K
from modlefork_amd import perturbate as P
from modlefork_amd import simulation as S
from modlefork_amd import workloads as W
from oracle import config as OC
from oracle import refcpu as R


def test_block_sums_known_answers_oracle():
    """contact_matrix_dense_test.cpp:134-164 ("CMatrix get w/ block"): the corner, off-diagonal
    and far-corner 5 x 5 blocks of a 100 x 100 matrix, with edge extension at both ends."""
    case = K.load()["cmatrix"]["CMatrix get w/ block"]
    band = K.block_band(case)
    for row, col, bs, want in case["block_checks"]:
        assert R.band_get_block(band, case["nrows"], case["ncols"], row, col, bs) == want


def test_block_sums_oracle_against_brute_force():
    """The oracle's block sum against a dense symmetric matrix built from the band, with every
    coordinate clamped to the matrix (Extend edge handling), on a narrow band so that blocks
    reach past it."""
    rng = np.random.default_rng(1)
    nrows, ncols = 7, 40
    band = rng.integers(0, 50, nrows * ncols + 1).astype(np.uint32)
    dense = np.zeros((ncols, ncols), np.int64)
    for j in range(ncols):
        for i in range(min(nrows, j + 1)):
            dense[j - i, j] = dense[j, j - i] = band[j * nrows + i]
    for _ in range(300):
        r, c = rng.integers(0, ncols, 2)
        bs = int(rng.choice([1, 3, 5, 9]))
        rr = np.clip(np.arange(r - bs // 2, r - bs // 2 + bs), 0, ncols - 1)
        cc = np.clip(np.arange(c - bs // 2, c - bs // 2 + bs), 0, ncols - 1)
        want = int(dense[np.ix_(rr, cc)].sum())
        assert R.band_get_block(band, nrows, ncols, int(r), int(c), bs) == want


def test_pdq_order_matches_the_oracle():
    rng = np.random.default_rng(2)
    for n in (0, 1, 5, 23, 24, 200, 3000):
        keys = rng.integers(0, max(1, n // 4), n).astype(np.uint64)
        assert P.pdq_order(keys).tolist() == list(R.pdq_order(keys.tolist()))


def test_barrier_import_order_matches_the_oracle():
    """Barriers whose BED records share starts and midpoints, given in file order: the
    chromosome's interval-tree order (two pdq_sorts by start around the range filter) and the
    simulation's position order agree with the oracle's field by field."""
    params, chroms, records = PC.build()
    c = OC.transform(PC.OPTIONS)
    for ch in chroms:
        recs = records[ch.name]
        for lo, hi in ((0, ch.size), (ch.size // 4, ch.size // 2)):
            imp = S.import_barriers(params, recs, lo, hi)
            rows = OC.import_barriers(recs, lo, hi, c)
            assert imp.bed_start.tolist() == [t[0] for t in rows]
            assert imp.bed_end.tolist() == [t[1] for t in rows]
            assert imp.pos.tolist() == [t[2] for t in rows]
            assert imp.dir.tolist() == [t[3] for t in rows]
            assert imp.stp_active.tolist() == [t[4] for t in rows]
            assert imp.stp_inactive.tolist() == [t[5] for t in rows]
            b = S.sort_barriers(imp)
            o = OC.sort_barriers(rows)
            for f in ("pos", "dir", "stp_active", "stp_inactive"):
                assert np.array_equal(getattr(b, f), getattr(o, f)), f
            assert np.all(np.diff(b.pos.astype(np.int64)) >= 0)


def test_generated_deletions_end_after_their_barrier():
    """generate_deletions (scheduler_common.cpp:51-68): [pos + 1 - size, pos + 1), shortened to
    [0, pos + 1) near the 5' end, one per barrier, as a tree ordered by start."""
    imp = S.ImportedBarriers(np.array([5, 20_000, 3_000, 20_000], np.uint64),
                             np.array([6, 20_001, 3_001, 20_001], np.uint64),
                             np.array([5, 20_000, 3_000, 20_000], np.uint64),
                             np.array([1, 2, 1, 2], np.uint8), np.zeros(4), np.zeros(4))
    t = P.generate_deletions(imp, 10_000)
    pairs = sorted(zip(t.start.tolist(), t.end.tolist()))
    assert pairs == [(0, 6), (0, 3_001), (10_001, 20_001), (10_001, 20_001)]
    assert np.all(np.diff(t.start.astype(np.int64)) >= 0)


def test_interval_tree_ranges():
    t = P.IntervalTree([10, 0, 30, 20], [15, 5, 35, 25])
    assert t.start.tolist() == [0, 10, 20, 30]
    assert t.equal_range(4, 21) == (0, 3)   # [0,5), [10,15), [20,25) overlap [4, 21)
    assert t.equal_range(5, 10) == (4, 4)   # half-open: nothing overlaps [5, 10)
    assert t.equal_range(34, 100) == (3, 4)


def test_windows_tile_the_chromosome():
    """The active windows advance by the band width D and cover the chromosome; each outer
    window holds its active window; the last active window ends at the chromosome end."""
    for end, D in ((6_000_000, 1_000_000), (6_500_000, 1_000_000), (2_000_000, 1_000_000)):
        ws = list(P.windows(end, D))
        covered = np.zeros(end // 1000, bool)
        for k, (a, b, c, d) in enumerate(ws):
            assert a <= c < d <= b <= max(end, 4 * D)
            covered[c // 1000:d // 1000] = True
            if k > 0:
                assert c == ws[k - 1][2] + D
        assert covered.all()
        assert min(ws[-1][3], end) == end


def test_plan_follows_the_reference_task_rules():
    """Tasks of run_perturbate (scheduler_perturbate.cpp:255-330) on the seeded genome: only the
    chromosome with both feature tracks, ids in order, cell ids counting per chromosome, each
    task's deletion overlapping its outer window, window ends clamped to the chromosome, LEFs
    and contact targets from the window's size."""
    params, chroms, _ = PC.build()
    tasks = P.plan(params, chroms)
    assert tasks and {t.chrom for t in tasks} == {"chrP"}
    assert [t.id for t in tasks] == list(range(len(tasks)))
    assert [t.cell_id for t in tasks] == list(range(len(tasks)))
    ch = chroms[0]
    dels = P.generate_deletions(ch.barriers, 10_000)
    per_window = {}
    for t in tasks:
        assert t.deletion_begin < t.window_end and t.deletion_begin + t.deletion_size > t.window_start
        assert t.window_end <= ch.end and t.active_window_end <= ch.end
        width = t.window_end - t.window_start
        assert t.num_lefs == round(width / 1e6 * 20.0)
        npix = -(-width // 5000) * (1_000_000 // 5000)
        assert t.num_target_contacts == max(1, round(0.05 * npix))
        assert t.num_target_epochs == P.USIZE_MAX
        per_window.setdefault((t.window_start, t.active_window_start), []).append(t)
    # every deletion overlapping a window gives one task of that window
    for (ws, aws), ts in per_window.items():
        we = ts[0].window_end
        want = [(int(a), int(b)) for a, b in zip(dels.start, dels.end) if a < we and b > ws]
        assert [(t.deletion_begin, t.deletion_begin + t.deletion_size) for t in ts] == want


def test_task_file_round_trip():
    params, chroms, _ = PC.build()
    tasks = P.plan(params, chroms)
    for t in tasks[::7]:
        line = P.task_line(t)
        assert len(line.split("\t")) == 15
        u = P.parse_task_line(line, chroms)
        assert u == t
    bad = P.task_line(tasks[0]).split("\t")
    bad[1] = "chrX"
    with pytest.raises(ValueError, match="Unable to find a chromosome"):
        P.parse_task_line("\t".join(bad), chroms)


def test_deletion_switches_barriers_off():
    """simulate_window's barrier configuration (scheduler_perturbate.cpp:463-472)."""
    params, chroms, _ = PC.build()
    tasks = P.plan(params, chroms)
    t = tasks[len(tasks) // 2]
    bars, nact = P.window_barriers(chroms[0], t)
    inside = (bars.pos >= t.deletion_begin) & (bars.pos < t.deletion_begin + t.deletion_size)
    assert inside.any()
    assert np.all(bars.stp_active[inside] == 0.0) and np.all(bars.stp_inactive[inside] == 1.0)
    assert nact == len(bars) - int(inside.sum())
    full = S.sort_barriers(chroms[0].barriers.subset(*t.barrier_range))
    assert np.array_equal(full.stp_active[~inside], bars.stp_active[~inside])


def test_binomial_test_two_sided():
    """The significance column is scipy.stats.binomtest's two-sided p-value, which
    stats::binomial_test restates (tests_impl.hpp:16-85; its own unit test compares with
    scipy); k == n / 2 gives 1."""
    assert P.binomial_test(5, 10) == 1.0
    assert abs(P.binomial_test(0, 10) - 2 * 0.5 ** 10) < 1e-15
    assert P.binomial_test(3, 10) == P.binomial_test(7, 10)
