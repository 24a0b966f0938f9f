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


def coarse_band(counts: np.ndarray, f: int = COARSE) -> np.ndarray:
    """[ND][N] counts of pixels (i, i + d) -> [ND // f][N // f] coarse counts of coarse pixels
    (I, I + D) (fine pixels whose row and column fall in coarse bins I and I + D)."""
    nd, n = counts.shape
    ncd, nc = nd // f, n // f
    out = np.zeros((ncd, nc), np.float64)
    d = np.arange(nd)[:, None]
    i = np.arange(n)[None, :]
    I = i // f
    D = (i + d) // f - I
    ok = (D < ncd) & (I < nc) & ((i + d) < nc * f)
    np.add.at(out, (D[ok], np.broadcast_to(I, counts.shape)[ok]), counts[ok])
    return out


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
                    seed: int = 0, diagonal_width: int = 3_000_000, chunk: int = 64):
    """Fast-mode simulation of each point (one parameter block per point); returns the [ND][nbins]
    count bands and the device seconds."""
    bands, dev_s = [], 0.0
    nbins = layout.size // Wn.BIN
    for c0 in range(0, len(points), chunk):
        part = points[c0:c0 + chunk]
        params, chroms, bars, plans = [], [], [], []
        for k, p in enumerate(part):
            prm = S.make_params(sweep_options(p, contact_density, seed, diagonal_width))
            ch = S.chrom_view(f"sweep{c0 + k}", layout.size)
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
    reliabilities: list     # the judge simulation's own reliability
    replaced: list
    fell_back: bool
    surrogate_s: float
    judge_s: float
    sim_rate: float         # simulated matrices per second (judge runs)
    surrogate_rate: float   # predicted matrices per second (inference only)
    sweep_rate: float       # matrices per second of the whole sweep (inference + judge)


def run_sweep(ctx, net, layout: Layout, points, judge_frac: float = 0.05, threshold: float = 0.95,
              seed: int = 3, ncells: int = 512, contact_density: float = 30.0, log=print):
    nbins = layout.size // Wn.BIN
    rng = np.random.default_rng(seed)
    nj = max(1, int(round(judge_frac * len(points))))
    judged = sorted(rng.choice(len(points), size=nj, replace=False).tolist())
    # surrogate: every point, windows batched across points
    xs = [point_inputs(layout, p) for p in points]
    nwin = xs[0].shape[0]
    t0 = time.perf_counter()
    pred = net.forward(np.concatenate(xs))
    t_sur = time.perf_counter() - t0
    rel = Wn.denormalise(pred).reshape(len(points), nwin, Wn.NDIAG, Wn.OUT_BINS)
    bands = [stitch(rel[k], nbins) for k in range(len(points))]
    # judge: each judged point simulated as two independent halves of its cells
    t0 = time.perf_counter()
    jp = [points[k] for k in judged]
    half_a, _ = simulate_points(ctx, layout, jp, ncells // 2, contact_density / 2, seed=seed)
    half_b, _ = simulate_points(ctx, layout, jp, ncells // 2, contact_density / 2, seed=seed + 1)
    t_judge = time.perf_counter() - t0
    corrs, raws, rels, replaced = [], [], [], []
    for k, a, b in zip(judged, half_a, half_b):
        c, r, q = judge(bands[k], a, b)
        corrs.append(c)
        raws.append(r)
        rels.append(q)
        if not c >= threshold:
            replaced.append(k)
            bands[k] = a + b
    fell_back = len(replaced) > 0.2 * len(judged)
    log(f"sweep: {len(points)} points, {nj} judged, corrected correlation "
        f"{np.nanmean(corrs):.4f} (raw {np.nanmean(raws):.4f}, judge reliability "
        f"{np.nanmean(rels):.4f}), {len(replaced)} replaced"
        f"{', surrogate rejected' if fell_back else ''}")
    return SweepResult(len(points), judged, corrs, raws, rels, replaced, fell_back, t_sur, t_judge,
                       nj / t_judge, len(points) / t_sur, len(points) / (t_sur + t_judge)), bands


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
    net = SG.Surrogate(ctx, SG.default_config(max_batch=32))
    SG.load_params(net, a.weights)
    layout = fixed_layout()
    points = sample_points(a.points)
    res, _ = run_sweep(ctx, net, layout, points, a.judge_frac, a.threshold, ncells=a.cells,
                       contact_density=a.density, log=log)
    # device-only inference rate: one batch of 32 windows
    nwin = len(window_starts(layout.size // Wn.BIN))
    ms = net.predict_timed(32, 10) / 10
    dev_rate = 32 / (ms / 1000.0) / nwin
    out = {"points": res.npoints, "judged": len(res.judged), "threshold": a.threshold,
           "corr_mean": float(np.nanmean(res.correlations)),
           "corr_min": float(np.nanmin(res.correlations)),
           "corr_p10": float(np.nanpercentile(res.correlations, 10)),
           "raw_corr_mean": float(np.nanmean(res.raw_correlations)),
           "judge_reliability_mean": float(np.nanmean(res.reliabilities)),
           "replaced": len(res.replaced), "surrogate_rejected": res.fell_back,
           "surrogate_wall_s": res.surrogate_s, "judge_wall_s": res.judge_s,
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
