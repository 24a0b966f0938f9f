"""Summary statistics of banded contact matrices used to compare fast mode with the oracle.

band layout: data[col * nrows + (col - row)] (ContactMatrixDense<u32>), nrows diagonals."""
import numpy as np


def as_diag_major(band, nrows, ncols):
    """(nrows, ncols) view of the band: entry [d, j] is pixel (j - d, j). No copy: bands of the
    large workloads hold 10^9 pixels, so every sum below accumulates in float64 without
    converting the band."""
    return np.asarray(band[:nrows * ncols]).reshape(ncols, nrows).T


def decay(band, nrows, ncols):
    """Mean contacts per pixel on each diagonal (the P(s) curve)."""
    m = as_diag_major(band, nrows, ncols)
    counts = np.maximum(ncols - np.arange(nrows), 1).astype(np.float64)
    return m.sum(axis=1, dtype=np.float64) / counts


def log_binned_decay(band, nrows, ncols, nbins=12):
    p = decay(band, nrows, ncols)
    edges = np.unique(np.geomspace(1, nrows, nbins + 1).astype(int))
    return np.array([p[a:b].mean() for a, b in zip(edges[:-1], edges[1:])]), edges


def coverage(band, nrows, ncols):
    """Contacts touching each bin (row and column sums of the band, diagonal counted once)."""
    m = as_diag_major(band, nrows, ncols)
    cov = m.sum(axis=0, dtype=np.float64)  # column j: pixels (j - d, j)
    for d in range(1, nrows):
        cov[: ncols - d] += m[d, d:]
    return cov


def block_sums(band, nrows, ncols, w=10):
    m = as_diag_major(band, nrows, ncols)
    r = (nrows // w) * w
    c = (ncols // w) * w
    return m[:r, :c].reshape(r // w, w, c // w, w).sum(axis=(1, 3), dtype=np.float64)


def pearson(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    a = a - a.mean()
    b = b - b.mean()
    return float((a * b).sum() / np.sqrt((a * a).sum() * (b * b).sum()))


def _stripe_means(m, nrows, ncols, bins, dirs, reach):
    """Per anchor bin b, the mean of pixels (b, b + k) (direction 1) or (b - k, b) (direction 2)
    for k in [1, reach), over the pixels inside the band; NaN where none is."""
    b = np.asarray(bins, np.int64)[:, None]
    k = np.arange(1, reach, dtype=np.int64)[None, :]
    dr = np.asarray(dirs)[:, None]
    j = np.where(dr == 1, b + k, b - k)  # REV-blocking barriers anchor loops to their right
    col = np.where(dr == 1, j, b)
    ok = (j >= 0) & (j < ncols) & (k < nrows)
    vals = np.where(ok, m[np.minimum(k, nrows - 1), np.clip(col, 0, ncols - 1)], 0).astype(np.float64)
    n = ok.sum(axis=1)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(n > 0, vals.sum(axis=1) / n, np.nan)


def barrier_stripe_enrichment(band, nrows, ncols, barrier_bins, dirs, reach=20):
    """Mean contacts along the stripe a barrier anchors (its row of pixels towards the side it
    blocks), relative to the same stripe around random bins."""
    m = as_diag_major(band, nrows, ncols)
    s = np.nanmean(_stripe_means(m, nrows, ncols, barrier_bins, dirs, reach))
    rng = np.random.default_rng(0)
    rnd = rng.integers(reach, ncols - reach, len(barrier_bins))
    ref = np.nanmean(_stripe_means(m, nrows, ncols, rnd, dirs, reach))
    return float(s / ref)


def decay_groups(band, nrows, ncols, min_contacts_ref=None, nbins=12):
    """P(s) over log-spaced diagonal groups; groups past the last one holding at least
    `min_contacts_ref` contacts of the reference band (an array of per-group contact sums) are
    pooled into one tail group. Returns (mean contacts per pixel per group, contacts per group)."""
    m = as_diag_major(band, nrows, ncols)
    per_diag = m.sum(axis=1, dtype=np.float64)
    edges = np.unique(np.geomspace(1, nrows, nbins + 1).astype(int))
    groups = list(zip(edges[:-1], edges[1:]))
    if min_contacts_ref is not None:
        keep = [g for g, n in zip(groups, min_contacts_ref) if n]
        groups = keep + [(keep[-1][1], nrows)] if keep and keep[-1][1] < nrows else keep
    sums = np.array([per_diag[a:b].sum() for a, b in groups])
    npix = np.array([sum(max(ncols - d, 0) for d in range(a, b)) for a, b in groups], float)
    return sums / npix, sums, groups


def barrier_dot_enrichment(band, nrows, ncols, barrier_bins, dirs, max_dist=40):
    """Mean observed/expected contacts at convergent barrier pairs: a REV-blocking barrier (dir 1)
    at bin i and a FWD-blocking one (dir 2) at bin j, 2 <= j - i < max_dist, where the expected
    value is the mean of diagonal j - i."""
    m = as_diag_major(band, nrows, ncols)
    p = decay(band, nrows, ncols)
    bins = np.asarray(barrier_bins, np.int64)
    dirs = np.asarray(dirs)
    rev = np.unique(bins[dirs == 1])
    fwd = np.unique(bins[dirs == 2])
    hi_d = min(nrows, max_dist)
    lo = np.searchsorted(fwd, rev + 2)
    hi = np.searchsorted(fwd, rev + hi_d)
    cnt = hi - lo
    if cnt.sum() == 0:
        return float("nan")
    ii = np.repeat(rev, cnt)
    jj = fwd[np.concatenate([np.arange(a, b) for a, b in zip(lo, hi) if b > a])]
    d = jj - ii
    ok = (jj < ncols) & (p[d] > 0)
    if not ok.any():
        return float("nan")
    return float(np.mean(m[d[ok], jj[ok]] / p[d[ok]]))
