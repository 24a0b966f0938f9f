"""Surrogate-accelerated parameter sweep with the plain simulation as judge and fallback.

A sweep point is (LEF density, processivity, occupancy scale) applied to one fixed barrier layout
of a 20 Mbp chromosome (5 kb bins). The surrogate predicts each point's contact band out to 100
diagonals by tiling the chromosome with overlapping windows; a seeded 5 % of the points are
also simulated (fast mode, 512 cells) and judged: a point whose mean per-diagonal correlation
with its simulation is below the threshold is replaced by the simulation, and when more than
a fifth of the judged points fail, the surrogate is not trusted for the sweep and every point
is simulated.

Agreement is measured at 25 kb (5 x 5 bins): counts are summed into coarse pixels, scaled to
mean 10 over the compared region and log1p-transformed; for each coarse diagonal 1 .. 19 on
which the judge's simulation holds at least MIN_COUNT contacts per coarse pixel, the Pearson
correlation runs along the diagonal. The judge simulation is itself a sample of 512 cells, so
its correlation with even the exact expected band is below 1; it runs as two independent
halves, their agreement gives its reliability (Spearman-Brown), and each diagonal's correlation
is divided by the square root of that reliability (Spearman's correction for attenuation) to
estimate the prediction's correlation with the expected band. The mean over the judged
diagonals is the score; the uncorrected correlation is reported beside it.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import simulation as S
from . import windows as Wn
from .workloads import BedRecord

COARSE = 5
MIN_COUNT = 100.0  # contacts per coarse pixel on a judged diagonal (Poisson noise <= 10 %)
MIN_RELIABILITY = 0.5  # judged diagonals: the correction divides by at most sqrt(2)
STRIDE = Wn.STRIDE
window_starts = Wn.window_starts


def _pearson(a, b):
    a = a - a.mean()
    b = b - b.mean()
    den = np.sqrt((a * a).sum() * (b * b).sum())
    return float((a * b).sum() / den) if den > 0 else 0.0


_COARSE_INDEX = {}


def _coarse_index(nd: int, n: int, f: int):
    """Flat coarse-pixel index D * (n // f) + I of every fine pixel (i, i + d) of a [nd][n] band,
    and the mask of the fine pixels that land in a coarse pixel of the [nd // f][n // f] band."""
    key = (nd, n, f)
    if key not in _COARSE_INDEX:
        ncd, nc = nd // f, n // f
        d = np.arange(nd)[:, None]
        i = np.arange(n)[None, :]
        I = i // f
        D = (i + d) // f - I
        ok = (D < ncd) & (I < nc) & ((i + d) < nc * f)
        _COARSE_INDEX[key] = ((D * nc + I)[ok].astype(np.int64), ok)
    return _COARSE_INDEX[key]


def coarse_band(counts: np.ndarray, f: int = COARSE) -> np.ndarray:
    """[ND][N] counts of pixels (i, i + d) -> [ND // f][N // f] coarse counts of coarse pixels
    (I, I + D) (fine pixels whose row and column fall in coarse bins I and I + D)."""
    nd, n = counts.shape
    ncd, nc = nd // f, n // f
    idx, ok = _coarse_index(nd, n, f)
    out = np.bincount(idx, weights=np.asarray(counts, np.float64)[ok], minlength=ncd * nc)
    return out.reshape(ncd, nc)


def diag_correlations(pred: np.ndarray, sim: np.ndarray, f: int = COARSE,
                      diagonals=None) -> np.ndarray:
    """Pearson correlation along each coarse diagonal D (default 1 .. ND/f - 1) of
    log1p(10 * coarse / mean coarse) between two [ND][N] count bands (pixels (i, i + d)); the mean
    is taken over the compared diagonals of each band."""
    a, b = coarse_band(pred, f), coarse_band(sim, f)
    nc = a.shape[1]
    ds = list(range(1, a.shape[0])) if diagonals is None else list(diagonals)
    if not ds:
        return np.zeros(0)
    xs = [a[D, :nc - D] for D in ds]  # coarse pixels (I, I + D) inside the band
    ys = [b[D, :nc - D] for D in ds]
    mx = np.mean(np.concatenate(xs))
    my = np.mean(np.concatenate(ys))
    return np.array([_pearson(np.log1p(10 * x / mx), np.log1p(10 * y / my)) for x, y in zip(xs, ys)])


def well_sampled(sim: np.ndarray, min_count: float = MIN_COUNT, f: int = COARSE):
    """Coarse diagonals on which `sim` holds at least min_count contacts per coarse pixel."""
    b = coarse_band(sim, f)
    nc = b.shape[1]
    return [D for D in range(1, b.shape[0]) if b[D, :nc - D].mean() >= min_count]


def diag_correlation(pred: np.ndarray, sim: np.ndarray, f: int = COARSE,
                     min_count: float = 0.0) -> float:
    """Mean of diag_correlations; with min_count, over the diagonals well sampled in `sim` only
    (nan when none is)."""
    ds = well_sampled(sim, min_count, f) if min_count > 0 else None
    c = diag_correlations(pred, sim, f, ds)
    return float(c.mean()) if len(c) else float("nan")


def judge(pred: np.ndarray, half_a: np.ndarray, half_b: np.ndarray, min_count: float = MIN_COUNT):
    """Agreement of a prediction with the expected band of a simulation run as two independent
    halves of its cells. Per well-sampled diagonal: r = corr(pred, a + b), the halves' agreement
    h = corr(a, b), the full simulation's reliability q = 2h / (1 + h) (Spearman-Brown), and
    r / sqrt(q), Spearman's correction for the simulation's own sampling noise (clipped to 1).
    Diagonals whose reliability is below MIN_RELIABILITY are left out. Returns (mean corrected,
    mean raw, mean reliability) over the judged diagonals."""
    full = half_a + half_b
    ds = well_sampled(full, min_count)
    if not ds:
        return float("nan"), float("nan"), float("nan")
    h = diag_correlations(half_a, half_b, diagonals=ds)
    q = 2 * h / (1 + h)
    # a diagonal the simulation cannot reproduce (reliability below MIN_RELIABILITY) says
    # nothing about the prediction; with none left the point is not judged (nan fails)
    keep = q >= MIN_RELIABILITY
    if not keep.any():
        return float("nan"), float("nan"), float(q.mean())
    ds = [D for D, k in zip(ds, keep) if k]
    r = diag_correlations(pred, full, diagonals=ds)
    q = q[keep]
    return float(np.mean(np.minimum(r / np.sqrt(q), 1.0))), float(r.mean()), float(q.mean())


def raw_fine(pred: np.ndarray, sim: np.ndarray) -> float:
    """Uncorrected agreement at the bands' own resolution (5 kb): the mean Pearson correlation
    of log1p(10 * count / mean) along every off-diagonal 1 .. ND - 1, whatever its counts."""
    return float(np.mean(diag_correlations(pred, sim, f=1)))


# ---- sweep points ----
@dataclass
class Layout:
    size: int
    records: list  # BedRecord with base occupancy as score


def fixed_layout(seed: int = 11, size: int = 20_000_000, nbar: int = 400) -> Layout:
    """The seeded 20 Mbp layout of the sweep: 400 barriers at uniform positions, random strand,
    base occupancy U(0.3, 0.95)."""
    rng = np.random.default_rng(seed)
    pos = np.sort(rng.choice(size - 2, size=nbar, replace=False) + 1)
    recs = [BedRecord("sweep", int(p), int(p) + 1, float(rng.uniform(0.3, 0.95)),
                      "+" if rng.random() < 0.5 else "-") for p in pos]
    return Layout(size, recs)


@dataclass
class Point:
    density: float
    processivity: float
    occ_scale: float


def sample_points(n: int, seed: int = 5) -> list[Point]:
    rng = np.random.default_rng(seed)
    return [Point(float(rng.uniform(5, 40)), float(rng.uniform(100e3, 600e3)),
                  float(rng.uniform(0.5, 1.0))) for _ in range(n)]


def scaled_records(layout: Layout, s: float):
    return [BedRecord(r.chrom, r.start, r.end, min(max(r.score * s, 1e-6), 1.0), r.strand)
            for r in layout.records]


def point_inputs(layout: Layout, p: Point) -> np.ndarray:
    """Inputs [nwin][IN_BINS][5] of the windows tiling the layout at point p."""
    nbins = layout.size // Wn.BIN
    recs = scaled_records(layout, p.occ_scale)
    return np.stack([Wn.track(recs, p.density, p.processivity, s - Wn.CTX_BINS, Wn.IN_BINS, nbins)
                     for s in window_starts(nbins)])


def points_inputs(layout: Layout, points) -> np.ndarray:
    """point_inputs for many points at once, [npoints * nwin][IN_BINS][6] (the same values):
    the count and direction channels come from the fixed layout, the occupancy channel from one
    scatter per point, and each window is a slice of the chromosome's tracks padded by the
    context on both sides."""
    nbins = layout.size // Wn.BIN
    starts = window_starts(nbins)
    npts = len(points)
    pad = Wn.CTX_BINS
    L = nbins + 2 * pad + Wn.IN_BINS
    bins = np.array([(r.start + r.end) // 2 // Wn.BIN for r in layout.records], np.int64)
    sign = np.array([1.0 if r.strand == "+" else -1.0 for r in layout.records], np.float32)
    base = np.array([r.score for r in layout.records], np.float64)
    cnt = np.zeros(L, np.float32)
    dirn = np.zeros(L, np.float32)
    inside = np.zeros(L, np.float32)
    np.add.at(cnt, bins + pad, 1.0)
    np.add.at(dirn, bins + pad, sign)
    inside[pad:pad + nbins] = 1.0
    scale = np.array([p.occ_scale for p in points], np.float64)
    occ_b = np.minimum(np.maximum(base[None, :] * scale[:, None], 1e-6), 1.0).astype(np.float32)
    occ = np.zeros((npts, L), np.float32)
    for b in range(len(bins)):  # bins may repeat: accumulate barrier by barrier, in record order
        occ[:, bins[b] + pad] += occ_b[:, b]
    s0 = np.array([Wn.scalar_channels(p.density, p.processivity)[0] for p in points], np.float32)
    s1 = np.array([Wn.scalar_channels(p.density, p.processivity)[1] for p in points], np.float32)
    x = np.empty((npts, len(starts), Wn.IN_BINS, Wn.IN_CH), np.float32)
    for k, s in enumerate(starts):
        sl = slice(s, s + Wn.IN_BINS)  # window bins s - CTX .. in padded coordinates
        x[:, k, :, 0] = cnt[sl]
        x[:, k, :, 1] = dirn[sl]
        x[:, k, :, 2] = occ[:, sl]
        x[:, k, :, 3] = inside[sl]
        x[:, k, :, 4] = s0[:, None]
        x[:, k, :, 5] = s1[:, None]
    return x.reshape(npts * len(starts), Wn.IN_BINS, Wn.IN_CH)


def stitch(pred_rel: np.ndarray, nbins: int) -> np.ndarray:
    """[ND][nbins] relative counts of a chromosome from its windows' predictions: pixel (i, i + d)
    comes from the last window starting at or before i whose output holds both ends."""
    starts = window_starts(nbins)
    out = np.zeros((Wn.NDIAG, nbins), np.float64)
    for k, s in enumerate(starts):
        lo = s  # rows [s_k, s_k+1): i + d < s_k+1 + 100 <= s_k + 400
        hi = nbins if k + 1 == len(starts) else starts[k + 1]
        for d in range(Wn.NDIAG):
            i = np.arange(lo, min(hi, nbins))
            ok = (i + d) < s + Wn.OUT_BINS
            i = i[ok & (i + d < nbins)]
            out[d, i] = pred_rel[k, d, i - s]
    return out


def sweep_options(p: Point, contact_density: float, seed: int, diagonal_width: int) -> dict:
    return {"bin_size": Wn.BIN, "diagonal_width": diagonal_width,
            "number_of_lefs_per_mbp": p.density, "avg_lef_processivity": int(round(p.processivity)),
            "target_contact_density": contact_density, "seed": seed}


def simulate_points(ctx, layout: Layout, points, ncells: int = 512, contact_density: float = 30.0,
                    seed: int = 0, diagonal_width: int = 3_000_000, chunk: int = 64,
                    stored_diagonals: int = 0, name_ids=None):
    """Fast-mode simulation of each point (one parameter block per point); returns the [ND][nbins]
    count bands and the device seconds. name_ids (one per point, optional) gives the index in a
    point's chromosome name ("sweep<i>", which enters its cells' seeds) instead of its position
    in `points`.

    stored_diagonals > 0 keeps only that many diagonals of each band on the device: the band is
    declared that wide and the contact density scaled by (full rows / stored rows), so the
    per-cell contact targets, and with them every cell's trajectory and every contact, are those
    of the full-width run; contacts beyond the stored diagonals are counted as missed. The bands
    returned (NDIAG diagonals) are the same, with less to clear and copy back."""
    bands, dev_s = [], 0.0
    nbins = layout.size // Wn.BIN
    if stored_diagonals:
        full_rows = min(-(-diagonal_width // Wn.BIN), nbins)
        rows = min(stored_diagonals, full_rows)
        contact_density = contact_density * full_rows / rows
        diagonal_width = rows * Wn.BIN
    for c0 in range(0, len(points), chunk):
        part = points[c0:c0 + chunk]
        params, chroms, bars, plans = [], [], [], []
        for k, p in enumerate(part):
            nid = c0 + k if name_ids is None else name_ids[c0 + k]
            prm = S.make_params(sweep_options(p, contact_density, seed, diagonal_width))
            ch = S.chrom_view(f"sweep{nid}", layout.size)
            params.append(prm)
            chroms.append(ch)
            bars.append(S.prepare_barriers(prm, scaled_records(layout, p.occ_scale), 0, layout.size))
            plans.append(S.plan_chrom(prm, ch, ncells))
        job = S.Job(ctx, params, chroms, bars, plans, L.MODE_FAST)
        st = job.run()
        dev_s += st.kernel_ms / 1000.0
        for k in range(len(part)):
            band, _ = job.band(k)
            bands.append(Wn.band_to_window(band, plans[k].nrows, 0, nbins))
        job.close()
    return bands, dev_s


@dataclass
class SweepResult:
    npoints: int
    judged: list
    correlations: list      # noise-corrected agreement of each judged point
    raw_correlations: list  # agreement with the (noisy) simulation itself
    raw_5kb: list           # uncorrected mean correlation at 5 kb over all 99 off-diagonals
    reliabilities: list     # the judge simulation's own reliability
    replaced: list
    fell_back: bool
    surrogate_s: float      # inference and stitching of every point, plus any input building
                            # the judge did not hide (wall)
    judge_s: float          # the judge's simulations (wall; inputs are built meanwhile)
    sweep_s: float          # the whole sweep, judge and fallback included (wall)
    sim_rate: float         # plain simulation, matrices per second, measured on its own
    surrogate_rate: float   # surrogate matrices per second (inputs, inference, stitching)
    sweep_rate: float       # matrices per second of the whole sweep
    slots: np.ndarray       # device band slot of each point (net.read_bands)


def plain_simulation_rate(ctx, layout: Layout, points, ncells: int = 512,
                          contact_density: float = 30.0, seed: int = 17) -> float:
    """Matrices per second of the plain method on its own: every point simulated with the judge's
    cells and contact density, the 100 diagonals the sweep delivers kept on the device."""
    t0 = time.perf_counter()
    simulate_points(ctx, layout, points, ncells, contact_density, seed=seed,
                    stored_diagonals=Wn.NDIAG, chunk=len(points))
    return len(points) / (time.perf_counter() - t0)


def run_sweep(ctx, net, layout: Layout, points, judge_frac: float = 0.05, threshold: float = 0.95,
              seed: int = 3, ncells: int = 512, contact_density: float = 30.0, log=print,
              rate_sample: int = 64, chunk_points: int = 256):
    """Predicts every point with the surrogate and judges a seeded judge_frac of them by
    simulation; a judged point whose score is below the threshold is replaced by its simulation,
    and more than a fifth failing rejects the surrogate for the sweep. The plain simulation's rate
    is then measured on its own on rate_sample judged points (0: not measured, sim_rate nan).

    Order of work: the judged points are predicted first; the judge then simulates them on its
    own context (HIP stream) in a host thread while the host builds every other point's inputs;
    after it, the device predicts the other points while a host thread scores the judged ones.
    (Running predictions beside the judge's persistent launches only makes them compete for the
    CUs.)

    Returns (SweepResult, replaced bands {point: band}). Every other point's band is the
    surrogate's and stays in the network's device memory: net.read_bands(result.slots[k], 1)."""
    t_start = time.perf_counter()
    nbins = layout.size // Wn.BIN
    rng = np.random.default_rng(seed)
    nj = max(1, int(round(judge_frac * len(points))))
    judged = sorted(rng.choice(len(points), size=nj, replace=False).tolist())
    jp = [points[k] for k in judged]
    js = set(judged)
    rest = [k for k in range(len(points)) if k not in js]
    order = judged + rest  # device band slot s holds point order[s]
    slots = np.empty(len(points), np.int64)
    slots[np.asarray(order)] = np.arange(len(points))
    starts = np.asarray(window_starts(nbins), np.uint32)
    nwin = len(starts)

    # 1. the judged points' predictions, read back for scoring
    t0 = time.perf_counter()
    net.reserve_bands(len(points), nbins)
    net.predict_chromosomes(points_inputs(layout, jp), 0, nwin, starts)
    pred_j = net.read_bands(0, nj)
    t_sur = time.perf_counter() - t0

    # 2. the judge on the device, the other points' inputs on the host
    out = {}

    def judge_worker():
        # each judged point simulated as two independent halves of its cells, both halves of
        # every judged point in one persistent launch (a half's cells take their seeds from its
        # own chromosome name, so the halves draw independent streams)
        jctx = S.Context(0)
        try:
            t1 = time.perf_counter()
            both, _ = simulate_points(jctx, layout, jp + jp, ncells // 2, contact_density / 2,
                                      seed=seed, stored_diagonals=Wn.NDIAG, chunk=2 * len(jp))
            out["a"], out["b"] = both[:len(jp)], both[len(jp):]
            out["t"] = time.perf_counter() - t1
        except BaseException as e:  # re-raised by the sweep thread
            out["err"] = e
        finally:
            jctx.close()

    th = threading.Thread(target=judge_worker, name="sweep-judge")
    th.start()
    t0 = time.perf_counter()
    try:
        xs = [(nj + c0, points_inputs(layout, [points[k] for k in rest[c0:c0 + chunk_points]]))
              for c0 in range(0, len(rest), chunk_points)]
    finally:
        t_inputs = time.perf_counter() - t0
        th.join()
    if "err" in out:
        raise out["err"]
    half_a, half_b, t_judge = out["a"], out["b"], out["t"]
    t_sur += max(0.0, t_inputs - t_judge)  # input building the judge did not hide

    # 3. the other points on the device, the judged points' scores on the host
    scores = {}

    def score_worker():
        try:
            scores["v"] = [judge(pred_j[j].astype(np.float64), a, b)
                           for j, (a, b) in enumerate(zip(half_a, half_b))]
            scores["raw5"] = [raw_fine(pred_j[j].astype(np.float64), a + b)
                              for j, (a, b) in enumerate(zip(half_a, half_b))]
        except BaseException as e:
            scores["err"] = e

    sc = threading.Thread(target=score_worker, name="sweep-score")
    sc.start()
    t0 = time.perf_counter()
    try:
        for slot, x in xs:
            net.predict_chromosomes(x, slot, nwin, starts)
    finally:
        t_sur += time.perf_counter() - t0
        sc.join()
    if "err" in scores:
        raise scores["err"]
    corrs, raws, rels, replaced = [], [], [], []
    bands = {}  # points whose band is the simulation's, not the surrogate's
    for k, (c, r, q), a, b in zip(judged, scores["v"], half_a, half_b):
        corrs.append(c)
        raws.append(r)
        rels.append(q)
        if not c >= threshold:
            replaced.append(k)
            bands[k] = a + b
    fell_back = len(replaced) > 0.2 * len(judged)
    if fell_back:
        # the surrogate is not trusted for this sweep: every point falls back to simulation
        sims, _ = simulate_points(ctx, layout, [points[k] for k in rest], ncells, contact_density,
                                  seed=seed, stored_diagonals=Wn.NDIAG)
        for k, b in zip(rest, sims):
            bands[k] = b
    t_sweep = time.perf_counter() - t_start
    log(f"sweep: {len(points)} points, {nj} judged, corrected correlation "
        f"{np.nanmean(corrs):.4f} (raw {np.nanmean(raws):.4f}, judge reliability "
        f"{np.nanmean(rels):.4f}), {len(replaced)} replaced"
        f"{', surrogate rejected' if fell_back else ''}; surrogate {t_sur:.1f} s, judge "
        f"{t_judge:.1f} s, sweep {t_sweep:.1f} s")
    sim_rate = (plain_simulation_rate(ctx, layout, jp[:rate_sample], ncells, contact_density)
                if rate_sample else float("nan"))
    return SweepResult(len(points), judged, corrs, raws, scores["raw5"], rels, replaced, fell_back,
                       t_sur, t_judge,
                       t_sweep, sim_rate, len(points) / t_sur, len(points) / t_sweep,
                       slots), bands


def main(argv=None):
    import argparse
    import json
    import os

    from . import surrogate as SG

    ap = argparse.ArgumentParser(description="surrogate-accelerated parameter sweep")
    ap.add_argument("--points", type=int, default=4096)
    ap.add_argument("--judge-frac", type=float, default=0.05)
    ap.add_argument("--threshold", type=float, default=0.95)
    ap.add_argument("--cells", type=int, default=512)
    ap.add_argument("--density", type=float, default=30.0, help="judge's target contact density")
    ap.add_argument("--weights", default=os.path.join(os.path.dirname(__file__), "data",
                                                      "surrogate_weights.npy"))
    ap.add_argument("--ceiling", type=int, default=0,
                    help="also simulate this many judged points a second time (other cell "
                         "streams) and report simulation-vs-simulation agreement")
    ap.add_argument("--summary", default=None)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    ctx = S.Context(0)
    layout = fixed_layout()
    nwin = len(window_starts(layout.size // Wn.BIN))
    net = SG.Surrogate(ctx, SG.default_config(max_batch=8 * nwin))
    SG.load_params(net, a.weights)
    points = sample_points(a.points)
    res, _ = run_sweep(ctx, net, layout, points, a.judge_frac, a.threshold, ncells=a.cells,
                       contact_density=a.density, log=log)
    # device-only inference rate: one batch of eight chromosomes' windows
    ms = net.predict_timed(8 * nwin, 10) / 10
    dev_rate = 8 * nwin / (ms / 1000.0) / nwin
    out = {"points": res.npoints, "judged": len(res.judged), "threshold": a.threshold,
           "corr_mean": float(np.nanmean(res.correlations)),
           "corr_min": float(np.nanmin(res.correlations)),
           "corr_p10": float(np.nanpercentile(res.correlations, 10)),
           "raw_corr_mean": float(np.nanmean(res.raw_correlations)),
           "raw_corr_5kb_all_diagonals_mean": float(np.nanmean(res.raw_5kb)),
           "judge_reliability_mean": float(np.nanmean(res.reliabilities)),
           "replaced": len(res.replaced), "surrogate_rejected": res.fell_back,
           "surrogate_wall_s": res.surrogate_s, "judge_wall_s": res.judge_s,
           "sweep_wall_s": res.sweep_s,
           "sim_matrices_per_s": res.sim_rate, "surrogate_matrices_per_s": res.surrogate_rate,
           "surrogate_device_matrices_per_s": dev_rate,
           "sweep_matrices_per_s": res.sweep_rate,
           "speedup_inference": res.surrogate_rate / res.sim_rate,
           "speedup_device_inference": dev_rate / res.sim_rate,
           "speedup_sweep": res.sweep_rate / res.sim_rate, "windows_per_point": nwin}
    if a.ceiling:
        pts = [points[k] for k in res.judged[:a.ceiling]]
        s1, _ = simulate_points(ctx, layout, pts, a.cells, a.density, seed=0)
        s2, _ = simulate_points(ctx, layout, pts, a.cells, a.density, seed=1)
        out["sim_vs_sim_corr"] = float(np.nanmean([diag_correlation(u, v, min_count=MIN_COUNT)
                                                    for u, v in zip(s1, s2)]))
    log(json.dumps(out))
    if a.summary:
        with open(a.summary, "w") as fh:
            json.dump(out, fh, indent=1)
    net.close()
    ctx.close()


if __name__ == "__main__":
    main()
