"""Host logic of the surrogate sweep: coarse-graining, window tiling and stitching, and the
judge's noise-corrected agreement, on synthetic bands whose answers are known."""
import numpy as np

from modlefork_amd import sweep as Sw
from modlefork_amd import windows as Wn


def test_coarse_band_conserves_complete_blocks():
    rng = np.random.default_rng(0)
    fine = rng.integers(0, 50, (100, 400)).astype(float)
    c = Sw.coarse_band(fine, 5)
    # coarse pixel (I, I + D) for D >= 1 sums the 25 fine pixels (5I + a, 5(I + D) + b)
    for I, D in [(0, 1), (7, 3), (40, 19), (60, 10)]:
        want = sum(fine[5 * (I + D) + b - (5 * I + a), 5 * I + a]
                   for a in range(5) for b in range(5))
        assert c[D, I] == want


def test_window_tiling_covers_every_band_pixel_once():
    for nbins in (400, 1000, 4000, 4321):
        starts = Wn.window_starts(nbins)
        assert starts[0] == 0 and starts[-1] == nbins - Wn.OUT_BINS
        rel = np.zeros((len(starts), Wn.NDIAG, Wn.OUT_BINS))
        for k in range(len(starts)):
            rel[k] = k + 1  # tag each window
        st = Sw.stitch(rel, nbins)
        d = np.arange(Wn.NDIAG)[:, None]
        i = np.arange(nbins)[None, :]
        inside = (i + d) < nbins
        assert (st[inside] > 0).all()
        # the window a pixel comes from holds both of its ends
        k = st.astype(int) - 1
        s = np.asarray(starts)[np.where(inside, k, 0)]
        assert ((i - s >= 0) & (i + d - s < Wn.OUT_BINS))[inside].all()


def test_judge_on_known_answers():
    rng = np.random.default_rng(1)
    expect = np.exp(rng.standard_normal((100, 2000)) * 0.5) * 100
    a = rng.poisson(expect).astype(float)
    b = rng.poisson(expect).astype(float)
    c, raw, q = Sw.judge(expect * 2, a, b)
    assert c > 0.99 and raw < c and 0.9 < q <= 1.0
    # an unrelated pattern scores near zero
    c, _, _ = Sw.judge(np.exp(rng.standard_normal((100, 2000)) * 0.5) * 100, a, b)
    assert abs(c) < 0.1
    # pure noise: the simulation cannot reproduce itself, so nothing is judged
    n = lambda: rng.poisson(100, (100, 2000)).astype(float)
    c, raw, q = Sw.judge(n(), n(), n())
    assert np.isnan(c) and q < Sw.MIN_RELIABILITY
    # too few contacts on every diagonal: not judged
    c, _, _ = Sw.judge(expect, a * 1e-3, b * 1e-3)
    assert np.isnan(c)


def test_normalise_round_trip():
    rng = np.random.default_rng(2)
    counts = rng.integers(0, 200, (3, Wn.NDIAG, Wn.OUT_BINS)).astype(float)
    y = Wn.normalise(counts)
    rel = Wn.denormalise(y)
    m = Wn.valid_mask()
    mean = np.where(m, counts, 0).sum(axis=(1, 2)) / m.sum()
    assert np.allclose(rel, counts / mean[:, None, None], rtol=1e-5)


def test_track_channels():
    from modlefork_amd.workloads import BedRecord
    recs = [BedRecord("c", 10_000, 10_001, 0.5, "+"), BedRecord("c", 12_000, 12_001, 0.25, "-"),
            BedRecord("c", 30_000, 30_001, 0.9, "+")]
    x = Wn.track(recs, 22.5, 350e3, -2, nbins=12, chrom_bins=8)
    # bins shifted by two: barrier bins 2, 2, 6 -> 4, 4, 8
    assert x[4, 0] == 2 and x[4, 1] == 0 and abs(x[4, 2] - 0.75) < 1e-6
    assert x[8, 0] == 1 and x[8, 1] == 1
    assert list(x[:, 3]) == [0, 0] + [1] * 8 + [0, 0]
    assert np.all(x[:, 4] == 0) and np.all(x[:, 5] == 0)


def test_batched_point_inputs_equal_the_per_point_tracks():
    layout = Sw.fixed_layout()
    pts = Sw.sample_points(5, seed=9)
    want = np.concatenate([Sw.point_inputs(layout, p) for p in pts])
    got = Sw.points_inputs(layout, pts)
    assert got.shape == want.shape
    assert np.array_equal(got, want)
