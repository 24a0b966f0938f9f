"""Generate the ziggurat layer tables used by the oracle's Boost.Random-equivalent
normal and exponential distributions.

TEST INFRASTRUCTURE: part of the CPU oracle. Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg use anything under oracle/.

Boost.Random 1.79 (`boost/random/normal_distribution.hpp`,
`boost/random/exponential_distribution.hpp`; the dependency pinned by the reference at
conanfile.py:17-40 and aliased at src/common/include/modle/common/random.hpp:36-59) draws
normals and exponentials with 128- and 256-layer ziggurats whose tables are literal
constants printed with 20 significant digits. Boost is absent here, so the tables are
rebuilt from their definition:

    x[1] = r, v = r f(r) + integral_r^inf f,  x[0] = v / f(r),
    x[i+1] = f^-1(f(x[i]) + v / x[i]),         y[i] = f(x[i]), y[0] = 0,
    x[N] = 0, y[N] = 1,

with f(x) = exp(-x^2/2), r = 3.4426198558966521214 (normal, N = 128) and
f(x) = exp(-x), r = 7.6971174701310497140 (exponential, N = 256), evaluated at 60
digits, rounded to 20 significant digits as Boost prints them, then to the nearest
double. The values are emitted as C99 hex-float literals so no decimal parsing is involved.
"""
import sys
import mpmath as mp

mp.mp.dps = 60


def build(r, f, finv, tail_area, n):
    r = mp.mpf(r)
    v = r * f(r) + tail_area(r)
    x = [mp.mpf(0)] * (n + 1)
    y = [mp.mpf(0)] * (n + 1)
    x[0] = v / f(r)
    x[1] = r
    y[1] = f(r)
    for i in range(1, n - 1):
        x[i + 1] = finv(y[i] + v / x[i])
        y[i + 1] = f(x[i + 1])
    x[n] = mp.mpf(0)
    y[n] = mp.mpf(1)
    return x, y


def as_double(v):
    # 20 significant digits, as the Boost headers print them, then nearest double
    return float(mp.nstr(v, 20, min_fixed=-30, max_fixed=30, strip_zeros=False)) if v != 0 else 0.0


def normal():
    f = lambda t: mp.e ** (-t * t / 2)
    finv = lambda yy: mp.sqrt(-2 * mp.log(yy))
    tail = lambda r: mp.sqrt(mp.pi / 2) * mp.erfc(r / mp.sqrt(2))
    return build("3.4426198558966521214", f, finv, tail, 128)


def exponential():
    f = lambda t: mp.e ** (-t)
    finv = lambda yy: -mp.log(yy)
    tail = lambda r: mp.e ** (-r)
    return build("7.6971174701310497140", f, finv, tail, 256)


def emit(name, vals):
    out = [f"static const double {name}[{len(vals)}] = {{"]
    for i in range(0, len(vals), 4):
        out.append("    " + ", ".join(as_double(v).hex() for v in vals[i:i + 4]) + ",")
    out.append("};")
    return "\n".join(out)


def main(path, prefix):
    nx, ny = normal()
    ex, ey = exponential()
    text = [
        "/* Generated by gen_zig_tables.py: Boost.Random-equivalent ziggurat layer tables. */",
        "#pragma once",
        emit(prefix + "normal_x", nx),
        emit(prefix + "normal_y", ny),
        emit(prefix + "exp_x", ex),
        emit(prefix + "exp_y", ey),
        "",
    ]
    with open(path, "w") as fh:
        fh.write("\n".join(text))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "zig_tables.h", sys.argv[2] if len(sys.argv) > 2 else "rc_zig_")
