"""Pins of the oracle's random stack: engine known answers, sampler distributions against
scipy, ziggurat tables against the digits Boost.Random prints, XXH3 against python-xxhash."""
import os
import random
import re

import numpy as np
import pytest
import xxhash
from scipy import stats

from oracle import refcpu as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
M64 = (1 << 64) - 1


def test_splitmix64_seed0_known_outputs():
    # SplitMix64 seeded with 0 (the generator xoshiro-cpp uses to expand a seed)
    g = R.rng(0)
    assert [hex(x) for x in g.s] == ["0xe220a8397b1dcdaf", "0x6e789e6aa1b965f4",
                                     "0x6c45d188009454f", "0xf88bb8a8724c81ec"]


def test_xoshiro256pp_hand_computed_steps():
    # state (1, 2, 3, 4): first output rotl(1 + 4, 23) + 1; second worked out by hand below
    g = R.RcRng()
    for i, v in enumerate((1, 2, 3, 4)):
        g.s[i] = v
    assert R.lib().rc_rng_next is not None
    R.lib().rc_rng_next.restype = __import__("ctypes").c_uint64
    R.lib().rc_rng_next.argtypes = [__import__("ctypes").POINTER(R.RcRng)]
    first = R.lib().rc_rng_next(g)
    assert first == (5 << 23) + 1
    # after one step: s0 = 1^(4^2) = 7, s3 = rotl(4^2, 45) = 6 << 45
    s0, s3 = 7, (6 << 45) & M64
    x = (s0 + s3) & M64
    assert R.lib().rc_rng_next(g) == ((((x << 23) | (x >> 41)) & M64) + s0) & M64


def draws(kind, n, a=0.0, b=0.0, c=0.0, seed=12345):
    return R.draw_sequence(seed, kind, n, a, b, c)


def test_normal_distribution():
    z = draws(1, 200_000, 4000.0, 200.0).view(np.float64)
    assert abs(z.mean() - 4000) < 5 * 200 / np.sqrt(len(z))
    assert abs(z.std() - 200) < 2.0
    assert stats.kstest((z - 4000) / 200, "norm").pvalue > 1e-3
    # the ziggurat tail (|z| > 3.44) is reached and has the right mass
    tail = np.mean(np.abs(z - 4000) / 200 > 3.4426198558966521)
    assert abs(tail - 2 * stats.norm.sf(3.4426198558966521)) < 5 * np.sqrt(5.8e-4 / len(z))


def test_exponential_distribution():
    x = draws(8, 200_000, 1.0).view(np.float64)
    assert stats.kstest(x, "expon").pvalue > 1e-3
    assert x.max() > 7.69711747  # beyond the first layer: the tail shift was used


def test_bernoulli_rate():
    for p in (0.02666, 0.5, 0.9):
        v = draws(2, 200_000, p)
        assert abs(v.mean() - p) < 5 * np.sqrt(p * (1 - p) / len(v))
    assert draws(2, 1000, 0.0).sum() == 0


def test_canonical_range():
    u = draws(3, 100_000).view(np.float64)
    assert u.min() >= 0 and u.max() < 1
    assert stats.kstest(u, "uniform").pvalue > 1e-3


@pytest.mark.parametrize("lo,hi", [(0, 399), (10, 13), (0, 19_999_999), (7, 7)])
def test_uniform_int(lo, hi):
    v = draws(4, 100_000, lo, hi).astype(np.int64)
    assert v.min() >= lo and v.max() <= hi
    if hi - lo < 1000 and hi > lo:
        counts = np.bincount(v - lo, minlength=hi - lo + 1)
        assert stats.chisquare(counts).pvalue > 1e-3
    if hi == lo:
        assert np.all(v == lo)


@pytest.mark.parametrize("mean", [2.13903743, 26.7379679, 150.0])
def test_poisson(mean):
    v = draws(5, 100_000, mean).astype(np.int64)
    ks = np.arange(0, int(mean + 8 * np.sqrt(mean) + 10))
    counts = np.bincount(np.clip(v, 0, ks[-1]), minlength=len(ks))[: len(ks)]
    expected = stats.poisson.pmf(ks, mean) * len(v)
    expected[-1] += stats.poisson.sf(ks[-1], mean) * len(v)
    keep = expected > 20
    obs = np.append(counts[keep], counts[~keep].sum())
    exp = np.append(expected[keep], expected[~keep].sum())
    assert stats.chisquare(obs, exp * obs.sum() / exp.sum()).pvalue > 1e-3


@pytest.mark.parametrize("t,p", [(64, 1 / 6), (320, 1 / 6), (40, 0.9), (800, 1 / 6)])
def test_binomial(t, p):
    v = draws(6, 100_000, t, p).astype(np.int64)
    assert v.min() >= 0 and v.max() <= t
    ks = np.arange(t + 1)
    counts = np.bincount(v, minlength=t + 1)
    expected = stats.binom.pmf(ks, t, p) * len(v)
    keep = expected > 20
    obs = np.append(counts[keep], counts[~keep].sum())
    exp = np.append(expected[keep], expected[~keep].sum())
    assert stats.chisquare(obs, exp * obs.sum() / exp.sum()).pvalue > 1e-3


def test_gev_matches_scipy_genextreme():
    # mu + sigma (1 - (-log u)^xi) / xi has CDF exp(-(1 - xi (x - mu) / sigma)^(1/xi)), which is
    # scipy's genextreme with shape c = xi
    x = draws(7, 100_000, 0.0, 5000.0, 0.001).view(np.float64)
    assert stats.kstest(x, stats.genextreme(c=0.001, loc=0.0, scale=5000.0).cdf).pvalue > 1e-3


BOOST_DIGITS = {
    "rc_zig_normal_x": {0: 3.7130862467403632609, 1: 3.4426198558966521214,
                        2: 3.2230849845786185446, 3: 3.0832288582142137009, 128: 0.0},
    "rc_zig_normal_y": {1: 0.0026696290839025035092, 2: 0.0055489952208164705392, 128: 1.0},
    "rc_zig_exp_x": {0: 8.6971174701310497140, 1: 7.6971174701310497140,
                     2: 6.9410336293772123602, 256: 0.0},
    "rc_zig_exp_y": {1: 0.00045413435384149675, 256: 1.0},
}


def parse_tables(path, prefix=""):
    text = open(path).read()
    out = {}
    for m in re.finditer(r"double (\w+)\[(\d+)\] = \{(.*?)\};", text, flags=re.S):
        vals = [float.fromhex(v.strip()) for v in m.group(3).split(",") if v.strip()]
        out[m.group(1)] = vals
    return out


def test_ziggurat_tables_match_boost_digits():
    t = parse_tables(os.path.join(ROOT, "oracle", "zig_tables.h"))
    assert len(t["rc_zig_normal_x"]) == 129 and len(t["rc_zig_exp_x"]) == 257
    for name, pins in BOOST_DIGITS.items():
        for i, v in pins.items():
            assert abs(t[name][i] - v) <= 2 * np.spacing(max(abs(v), 1e-300)), (name, i)
    # layers have equal area v: x[i] * (y[i+1] - y[i]) is constant (normal f = exp(-x^2/2))
    x, y = np.array(t["rc_zig_normal_x"]), np.array(t["rc_zig_normal_y"])
    areas = x[1:127] * (y[2:128] - y[1:127])
    assert np.ptp(areas) / areas.mean() < 1e-12
    assert np.allclose(y[1:128], np.exp(-x[1:128] ** 2 / 2), rtol=1e-15)


def test_device_tables_equal_oracle_tables():
    o = parse_tables(os.path.join(ROOT, "oracle", "zig_tables.h"))
    d = parse_tables(os.path.join(ROOT, "modlefork_amd", "csrc", "rng_tables.h"))
    assert o["rc_zig_normal_x"] == d["kZigNormalX"]
    assert o["rc_zig_normal_y"] == d["kZigNormalY"]
    assert o["rc_zig_exp_x"] == d["kZigExpX"]
    assert o["rc_zig_exp_y"] == d["kZigExpY"]


def test_oracle_xxh3_against_python_xxhash():
    rnd = random.Random(7)
    for n in list(range(0, 241)):
        data = bytes(rnd.getrandbits(8) for _ in range(n))
        seed = rnd.getrandbits(64)
        assert R.lib().rc_xxh3_64_withseed(data, n, seed) == xxhash.xxh3_64_intdigest(data, seed=seed)
