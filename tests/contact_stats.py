"""Summary statistics of banded contact matrices used to compare fast mode with the oracle.

band layout: data[col * nrows + (col - row)] (ContactMatrixDense<u32>), nrows diagonals."""
import numpy as np


def as_diag_major(band, nrows, ncols):
    """(nrows, ncols) array: entry [d, j] is pixel (j - d, j)."""
    return np.asarray(band[:nrows * ncols], dtype=np.float64).reshape(ncols, nrows).T


def decay(band, nrows, ncols):
    """Mean contacts per pixel on each diagonal (the P(s) curve)."""
    m = as_diag_major(band, nrows, ncols)
    counts = np.array([max(ncols - d, 1) for d in range(nrows)], dtype=np.float64)
    return m.sum(axis=1) / counts


def log_binned_decay(band, nrows, ncols, nbins=12):
    p = decay(band, nrows, ncols)
    edges = np.unique(np.geomspace(1, nrows, nbins + 1).astype(int))
    return np.array([p[a:b].mean() for a, b in zip(edges[:-1], edges[1:])]), edges


def coverage(band, nrows, ncols):
    """Contacts touching each bin (row and column sums of the band, diagonal counted once)."""
    m = as_diag_major(band, nrows, ncols)
    cov = m.sum(axis=0).copy()  # column j: pixels (j - d, j)
    for d in range(1, nrows):
        cov[: ncols - d] += m[d, d:]
    return cov


def block_sums(band, nrows, ncols, w=10):
    m = as_diag_major(band, nrows, ncols)
    r = (nrows // w) * w
    c = (ncols // w) * w
    return m[:r, :c].reshape(r // w, w, c // w, w).sum(axis=(1, 3))


def pearson(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    a = a - a.mean()
    b = b - b.mean()
    return float((a * b).sum() / np.sqrt((a * a).sum() * (b * b).sum()))


def barrier_stripe_enrichment(band, nrows, ncols, barrier_bins, dirs, reach=20):
    """Mean contacts along the stripe a barrier anchors (its row of pixels towards the side it
    blocks), relative to the same stripe around random bins."""
    m = as_diag_major(band, nrows, ncols)

    def pixel(i, j):
        if i > j:
            i, j = j, i
        d = j - i
        return m[d, j] if d < nrows and 0 <= i and j < ncols else np.nan

    def stripe(b, dr):
        vals = []
        for k in range(1, reach):
            j = b + k if dr == 1 else b - k  # REV-blocking barriers anchor loops to their right
            if 0 <= j < ncols:
                vals.append(pixel(b, j))
        return np.nanmean(vals) if vals else np.nan

    s = np.nanmean([stripe(b, d) for b, d in zip(barrier_bins, dirs)])
    rng = np.random.default_rng(0)
    ref = np.nanmean([stripe(b, d) for b, d in
                      zip(rng.integers(reach, ncols - reach, len(barrier_bins)), dirs)])
    return float(s / ref)


def decay_groups(band, nrows, ncols, min_contacts_ref=None, nbins=12):
    """P(s) over log-spaced diagonal groups; groups past the last one holding at least
    `min_contacts_ref` contacts of the reference band (an array of per-group contact sums) are
    pooled into one tail group. Returns (mean contacts per pixel per group, contacts per group)."""
    m = as_diag_major(band, nrows, ncols)
    edges = np.unique(np.geomspace(1, nrows, nbins + 1).astype(int))
    groups = list(zip(edges[:-1], edges[1:]))
    if min_contacts_ref is not None:
        keep = [g for g, n in zip(groups, min_contacts_ref) if n]
        groups = keep + [(keep[-1][1], nrows)] if keep and keep[-1][1] < nrows else keep
    sums = np.array([m[a:b].sum() for a, b in groups])
    npix = np.array([sum(max(ncols - d, 0) for d in range(a, b)) for a, b in groups], float)
    return sums / npix, sums, groups


def barrier_dot_enrichment(band, nrows, ncols, barrier_bins, dirs, max_dist=40):
    """Mean observed/expected contacts at convergent barrier pairs: a REV-blocking barrier (dir 1)
    at bin i and a FWD-blocking one (dir 2) at bin j, 2 <= j - i < max_dist, where the expected
    value is the mean of diagonal j - i."""
    m = as_diag_major(band, nrows, ncols)
    p = decay(band, nrows, ncols)
    rev = sorted({int(b) for b, d in zip(barrier_bins, dirs) if d == 1})
    fwd = sorted({int(b) for b, d in zip(barrier_bins, dirs) if d == 2})
    ratios = []
    for i in rev:
        for j in fwd:
            d = j - i
            if 2 <= d < min(nrows, max_dist) and j < ncols and p[d] > 0:
                ratios.append(m[d, j] / p[d])
    return float(np.mean(ratios)) if ratios else float("nan")
