"""Windows for the surrogate: sampled layouts, their input tracks and simulated targets.

A window is a 2 Mbp stretch (400 bins of 5 kb) whose contact band out to 100 diagonals (500 kb)
the surrogate predicts. Its input covers 500 kb of context on each side (600 bins). Training
windows tile long simulated chromosomes (20 Mbp, one LEF density and processivity each, barrier
density varying from one 2 Mbp segment to the next) exactly as a sweep tiles its chromosome, so
the network learns windows inside a chromosome and at its ends. Training targets are scaled by
the mean of the whole chromosome's band, so a window keeps its level relative to the rest.

Input channels per bin: barrier count, net blocking direction (+1 per REV-blocking, -1 per
FWD-blocking barrier), summed occupancy, 1 inside the chromosome (0 in the padding beyond its
ends), then LEF density and processivity broadcast along the window ((d - 22.5) / 10 and
(p - 350 kb) / 150 kb). Target: log1p(10 * count / mean count over the window's valid pixels)
for pixel (i, i + d), entry [d][i].
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import simulation as S
from .workloads import BedRecord

BIN = 5000
OUT_BINS = 400
CTX_BINS = 100
IN_BINS = OUT_BINS + 2 * CTX_BINS
NDIAG = 100
IN_CH = 6
CHROM_BINS = 4000  # training chromosomes: 20 Mbp
TARGET_SCALE = 10.0


@dataclass
class WindowSpec:
    density: float       # LEFs per Mbp
    processivity: float  # bp
    records: list        # BedRecord barriers on [0, IN_BINS * BIN)


def sample_specs(n: int, seed: int, max_barriers: int = 60) -> list[WindowSpec]:
    """Window layouts: 0-60 barriers per 2 Mbp at uniform positions (the same rate over the
    context), random strand, occupancy U(0.3, 0.95); LEF density U(5, 40)/Mbp; processivity
    U(100, 600) kb."""
    rng = np.random.default_rng(seed)
    span = IN_BINS * BIN
    out = []
    for _ in range(n):
        k_out = int(rng.integers(0, max_barriers + 1))
        k = int(rng.poisson(k_out * IN_BINS / OUT_BINS)) if k_out else 0
        pos = np.sort(rng.choice(span - 2, size=k, replace=False) + 1) if k else np.zeros(0, int)
        recs = [BedRecord("window", int(p), int(p) + 1, float(rng.uniform(0.3, 0.95)),
                          "+" if rng.random() < 0.5 else "-") for p in pos]
        out.append(WindowSpec(float(rng.uniform(5, 40)), float(rng.uniform(100e3, 600e3)), recs))
    return out


def scalar_channels(density: float, processivity: float):
    return (density - 22.5) / 10.0, (processivity - 350e3) / 150e3


def track(records, density: float, processivity: float, first_bin: int, nbins: int = IN_BINS,
          chrom_bins: int | None = None):
    """Input channels [nbins][6] of bins first_bin .. first_bin + nbins of a layout whose
    chromosome has chrom_bins bins (None: the window lies inside it)."""
    x = np.zeros((nbins, IN_CH), np.float32)
    for r in records:
        b = (r.start + r.end) // 2 // BIN - first_bin
        if 0 <= b < nbins:
            x[b, 0] += 1.0
            x[b, 1] += 1.0 if r.strand == "+" else -1.0
            x[b, 2] += float(r.score)
    g = first_bin + np.arange(nbins)
    x[:, 3] = 1.0 if chrom_bins is None else ((g >= 0) & (g < chrom_bins)).astype(np.float32)
    s0, s1 = scalar_channels(density, processivity)
    x[:, 4] = s0
    x[:, 5] = s1
    return x


def window_options(spec: WindowSpec, contact_density: float, seed: int) -> dict:
    return {"bin_size": BIN, "diagonal_width": NDIAG * BIN, "number_of_lefs_per_mbp": spec.density,
            "avg_lef_processivity": int(round(spec.processivity)),
            "target_contact_density": contact_density, "seed": seed}


def band_to_window(band: np.ndarray, nrows: int, first_bin: int, nbins: int = OUT_BINS):
    """Counts [NDIAG][nbins] of pixels (i, i + d) for i in the window (ContactMatrixDense layout
    data[col * nrows + d]); pixels beyond the band or the chromosome are 0."""
    ncols = (len(band) - 1) // nrows
    out = np.zeros((NDIAG, nbins), np.float32)
    m = band[:nrows * ncols].reshape(ncols, nrows)
    for d in range(min(NDIAG, nrows)):
        i = np.arange(nbins)
        col = first_bin + i + d
        ok = (col < ncols) & (first_bin + i >= 0)
        out[d, ok] = m[col[ok], d]
    return out


def valid_mask(nbins: int = OUT_BINS):
    d = np.arange(NDIAG)[:, None]
    i = np.arange(nbins)[None, :]
    return (i + d) < nbins


def normalise(counts: np.ndarray, mean=None) -> np.ndarray:
    """log1p(10 * count / mean) per window ([..., NDIAG, OUT_BINS]); mean defaults to the
    window's own mean count over its valid pixels (a scalar or one value per window)."""
    if mean is None:
        m = valid_mask(counts.shape[-1])
        mean = np.where(m, counts, 0).sum(axis=(-2, -1), keepdims=True) / m.sum()
    else:
        mean = np.asarray(mean, np.float64).reshape((-1,) + (1,) * 2 if np.ndim(mean) else ())
    return np.log1p(TARGET_SCALE * counts / np.maximum(mean, 1e-12)).astype(np.float32)


def band_mean(counts: np.ndarray) -> float:
    """Mean count over the pixels (i, i + d < N) of a [NDIAG][N] band."""
    nd, n = counts.shape
    m = (np.arange(n)[None, :] + np.arange(nd)[:, None]) < n
    return float(np.where(m, counts, 0).sum() / m.sum())


def denormalise(y: np.ndarray) -> np.ndarray:
    """Relative counts (mean 1 over the valid pixels, up to the log's distortion)."""
    return np.expm1(y) / TARGET_SCALE


def simulate_windows(ctx: S.Context, specs, ncells: int = 256, contact_density: float = 10.0,
                     seed: int = 0, chunk: int = 1024, threads: int = 0):
    """Simulates every window (fast mode, one job per chunk of windows, one parameter block per
    window) and returns (x [n][IN_BINS][5], counts [n][NDIAG][OUT_BINS], kernel seconds)."""
    xs, ys = [], []
    kernel_s = 0.0
    for c0 in range(0, len(specs), chunk):
        part = specs[c0:c0 + chunk]
        params, chroms, bars, plans = [], [], [], []
        for k, sp in enumerate(part):
            p = S.make_params(window_options(sp, contact_density, seed))
            ch = S.chrom_view(f"window{c0 + k}", IN_BINS * BIN)
            params.append(p)
            chroms.append(ch)
            bars.append(S.prepare_barriers(p, sp.records, 0, IN_BINS * BIN))
            plans.append(S.plan_chrom(p, ch, ncells))
            xs.append(track(sp.records, sp.density, sp.processivity, 0, IN_BINS, IN_BINS))
        job = S.Job(ctx, params, chroms, bars, plans, L.MODE_FAST)
        if threads:
            job.set_tuning(threads)
        st = job.run()
        kernel_s += st.kernel_ms / 1000.0
        for k in range(len(part)):
            band, _ = job.band(k)
            ys.append(band_to_window(band, plans[k].nrows, CTX_BINS))
        job.close()
    return np.stack(xs), np.stack(ys), kernel_s


# ---------------------------------------------------------------------------------------------
# training windows tiled over long chromosomes
# ---------------------------------------------------------------------------------------------
STRIDE = 300  # window starts; every pixel (i, i + d < i + 100) lies inside one output window


def window_starts(nbins: int):
    last = max(0, nbins - OUT_BINS)
    starts = list(range(0, last + 1, STRIDE))
    if starts[-1] != last:
        starts.append(last)
    return starts


@dataclass
class ChromSpec:
    density: float
    processivity: float
    records: list  # BedRecord barriers on [0, CHROM_BINS * BIN)


def sample_chromosomes(n: int, seed: int, max_barriers: int = 60) -> list[ChromSpec]:
    """20 Mbp layouts: per 2 Mbp segment 0-60 barriers at uniform positions, random strand,
    occupancy U(0.3, 0.95) times a per-chromosome scale U(0.5, 1); LEF density U(5, 40)/Mbp and
    processivity U(100, 600) kb per chromosome."""
    rng = np.random.default_rng(seed)
    seg = OUT_BINS * BIN
    out = []
    for _ in range(n):
        scale = float(rng.uniform(0.5, 1.0))
        recs = []
        for s0 in range(0, CHROM_BINS * BIN, seg):
            k = int(rng.integers(0, max_barriers + 1))
            pos = np.sort(rng.choice(seg - 2, size=k, replace=False) + 1 + s0) if k else []
            recs += [BedRecord("train", int(p), int(p) + 1, float(rng.uniform(0.3, 0.95)) * scale,
                               "+" if rng.random() < 0.5 else "-") for p in pos]
        out.append(ChromSpec(float(rng.uniform(5, 40)), float(rng.uniform(100e3, 600e3)), recs))
    return out


def chrom_windows(spec: ChromSpec):
    """Inputs [nwin][IN_BINS][6] of the windows tiling one chromosome."""
    return np.stack([track(spec.records, spec.density, spec.processivity, s - CTX_BINS, IN_BINS,
                           CHROM_BINS) for s in window_starts(CHROM_BINS)])


def chrom_options(spec: ChromSpec, contact_density: float, seed: int) -> dict:
    return {"bin_size": BIN, "diagonal_width": NDIAG * BIN, "number_of_lefs_per_mbp": spec.density,
            "avg_lef_processivity": int(round(spec.processivity)),
            "target_contact_density": contact_density, "seed": seed}


def simulate_chromosomes(ctx: S.Context, specs, ncells: int = 512, contact_density: float = 30.0,
                         seed: int = 0, chunk: int = 64, log=None):
    """Simulates each chromosome (fast mode, one parameter block per chromosome) and cuts it into
    windows: (x [n * nwin][IN_BINS][6], counts [n * nwin][NDIAG][OUT_BINS], the band mean of each
    window's chromosome [n * nwin], kernel seconds)."""
    xs, ys, means = [], [], []
    kernel_s = 0.0
    starts = window_starts(CHROM_BINS)
    for c0 in range(0, len(specs), chunk):
        part = specs[c0:c0 + chunk]
        params, chroms, bars, plans = [], [], [], []
        for k, sp in enumerate(part):
            p = S.make_params(chrom_options(sp, contact_density, seed))
            ch = S.chrom_view(f"train{c0 + k}", CHROM_BINS * BIN)
            params.append(p)
            chroms.append(ch)
            bars.append(S.prepare_barriers(p, sp.records, 0, CHROM_BINS * BIN))
            plans.append(S.plan_chrom(p, ch, ncells))
            xs.append(chrom_windows(sp))
        job = S.Job(ctx, params, chroms, bars, plans, L.MODE_FAST)
        st = job.run()
        kernel_s += st.kernel_ms / 1000.0
        if log:
            log(f"  simulated chromosomes {c0 + len(part)}/{len(specs)} ({kernel_s:.1f} s in kernel)")
        for k in range(len(part)):
            band, _ = job.band(k)
            full = band_to_window(band, plans[k].nrows, 0, CHROM_BINS)
            ys.extend(full[:, s:s + OUT_BINS] for s in starts)
            means.extend([band_mean(full)] * len(starts))
        job.close()
    return np.concatenate(xs), np.stack(ys), np.array(means), kernel_s
