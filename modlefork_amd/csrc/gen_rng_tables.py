"""Writes rng_tables.h: the 128-layer normal and 256-layer exponential ziggurat tables used by
the deterministic-mode samplers, matching the constants of Boost.Random 1.79
(boost/random/normal_distribution.hpp, exponential_distribution.hpp).

Each table is built from its definition (x[1] = r, x[0] = v / f(r),
x[i+1] = f^-1(f(x[i]) + v / x[i]), y[i] = f(x[i]), x[N] = 0, y[N] = 1) at 60 digits with
mpmath, rounded to the 20 significant digits Boost prints, then to double, and written as
hex-float literals so the compiler does no decimal rounding.
"""
import sys

import mpmath as mp

mp.mp.dps = 60


def layers(r, f, finv, tail, n):
    r = mp.mpf(r)
    v = r * f(r) + tail(r)
    x = [mp.mpf(0)] * (n + 1)
    y = [mp.mpf(0)] * (n + 1)
    x[0], x[1], y[1] = v / f(r), r, f(r)
    for i in range(1, n - 1):
        x[i + 1] = finv(y[i] + v / x[i])
        y[i + 1] = f(x[i + 1])
    y[n] = mp.mpf(1)
    return x, y


def to_double(v):
    if v == 0:
        return 0.0
    return float(mp.nstr(v, 20, min_fixed=-30, max_fixed=30, strip_zeros=False))


def table(name, vals):
    rows = [", ".join(to_double(v).hex() for v in vals[i:i + 4]) for i in range(0, len(vals), 4)]
    return (f"__device__ __constant__ double {name}[{len(vals)}] = {{\n    "
            + ",\n    ".join(rows) + "};\n")


def main(out):
    nx, ny = layers("3.4426198558966521214", lambda t: mp.e ** (-t * t / 2),
                    lambda yy: mp.sqrt(-2 * mp.log(yy)),
                    lambda r: mp.sqrt(mp.pi / 2) * mp.erfc(r / mp.sqrt(2)), 128)
    ex, ey = layers("7.6971174701310497140", lambda t: mp.e ** (-t), lambda yy: -mp.log(yy),
                    lambda r: mp.e ** (-r), 256)
    with open(out, "w") as fh:
        fh.write("// Generated by gen_rng_tables.py. Ziggurat layer tables (normal: 128 layers,\n"
                 "// exponential: 256 layers) equal to Boost.Random 1.79's constants.\n#pragma once\n")
        fh.write(table("kZigNormalX", nx))
        fh.write(table("kZigNormalY", ny))
        fh.write(table("kZigExpX", ex))
        fh.write(table("kZigExpY", ey))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "rng_tables.h")
