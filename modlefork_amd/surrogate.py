"""Learned surrogate of the simulation over the C-ABI (mgpu_sg_*).

The network maps a window's barrier track (position/direction/occupancy channels) and two
scalars (LEF density, processivity) to the window's log1p contact band; forward, backward and
AdamW run as the library's HIP kernels (BF16 on MFMA, FP32 accumulation). It has no counterpart
in the reference. Arrays are fp32 numpy: x [n][in_len][in_ch], y [n][ndiag][out_len].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

FP = C.POINTER(C.c_float)


def _f(a):
    return a.ctypes.data_as(FP)


def default_config(**over) -> L.MgpuSgConfig:
    c = L.MgpuSgConfig()
    L.check(L.lib().mgpu_sg_config_default(C.byref(c)))
    for k, v in over.items():
        if k == "enc_dilation":
            for i in range(16):
                c.enc_dilation[i] = v[i] if i < len(v) else 0
        else:
            setattr(c, k, v)
    return c


class Surrogate:
    def __init__(self, ctx, cfg: L.MgpuSgConfig | None = None):
        self.ctx = ctx
        self.cfg = cfg or default_config()
        self.ptr = C.c_void_p()
        rc = L.lib().mgpu_sg_create(ctx.ptr, C.byref(self.cfg), C.byref(self.ptr))
        if rc != 0:
            raise RuntimeError(f"mgpu_sg_create failed ({rc}): "
                               f"{L.lib().mgpu_sg_last_error(None).decode()}")
        self.nparams = int(L.lib().mgpu_sg_num_params(self.ptr))

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"surrogate call failed ({rc}): "
                               f"{L.lib().mgpu_sg_last_error(self.ptr).decode()}")

    @property
    def out_shape(self):
        return (int(self.cfg.ndiag), int(self.cfg.out_len))

    def params(self) -> np.ndarray:
        out = np.zeros(self.nparams, np.float32)
        self._check(L.lib().mgpu_sg_get_params(self.ptr, _f(out)))
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        self._check(L.lib().mgpu_sg_set_params(self.ptr, _f(p)))

    def _check_x(self, x):
        want = (int(self.cfg.in_len), int(self.cfg.in_ch))
        if x.ndim != 3 or tuple(x.shape[1:]) != want:
            raise ValueError(f"inputs must be [n][{want[0]}][{want[1]}], got {list(x.shape)}")

    def _check_y(self, y, n):
        if tuple(y.shape) != (n,) + self.out_shape:
            raise ValueError(f"targets must be [{n}][{self.out_shape[0]}][{self.out_shape[1]}], "
                             f"got {list(y.shape)}")

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        self._check_x(x)
        n = x.shape[0]
        y = np.zeros((n,) + self.out_shape, np.float32)
        self._check(L.lib().mgpu_sg_forward(self.ptr, _f(x), n, _f(y)))
        return y

    def reserve_bands(self, npoints: int, nbins: int):
        """Device memory for npoints whole-chromosome bands of nbins bins."""
        self._check(L.lib().mgpu_sg_reserve_bands(self.ptr, int(npoints), int(nbins)))
        self._bands_nbins = int(nbins)

    def predict_chromosomes(self, x: np.ndarray, first_point: int, nwin: int, starts) -> None:
        """Predicts chromosomes first_point .. from their windows' inputs x
        ([npoints * nwin][in_len][in_ch], window k of a point starting at output bin starts[k]);
        the stitched relative-count bands stay in device memory (read_bands copies them)."""
        x = np.ascontiguousarray(x, np.float32)
        self._check_x(x)
        if x.shape[0] % nwin:
            raise ValueError("inputs must hold whole chromosomes (a multiple of nwin windows)")
        st = np.ascontiguousarray(starts, np.uint32)
        if len(st) != nwin:
            raise ValueError("one start per window expected")
        self._check(L.lib().mgpu_sg_predict_chromosomes(
            self.ptr, _f(x), int(first_point), x.shape[0] // nwin, int(nwin),
            st.ctypes.data_as(C.POINTER(C.c_uint32))))

    def read_bands(self, first_point: int, npoints: int) -> np.ndarray:
        out = np.zeros((npoints, int(self.cfg.ndiag), self._bands_nbins), np.float32)
        self._check(L.lib().mgpu_sg_read_bands(self.ptr, int(first_point), int(npoints), _f(out)))
        return out

    def gradients(self, x: np.ndarray, y: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        self._check_x(x)
        self._check_y(y, x.shape[0])
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_float()
        self._check(L.lib().mgpu_sg_gradients(self.ptr, _f(x), _f(y), x.shape[0], C.byref(loss),
                                              _f(g)))
        return float(loss.value), g

    def adamw_step(self, g: np.ndarray, lr: float, weight_decay: float):
        g = np.ascontiguousarray(g, np.float32)
        self._check(L.lib().mgpu_sg_adamw_step(self.ptr, _f(g), lr, weight_decay))

    def set_data(self, x: np.ndarray, y: np.ndarray):
        self._x = np.ascontiguousarray(x, np.float32)
        self._y = np.ascontiguousarray(y, np.float32)
        self._check_x(self._x)
        self._check_y(self._y, self._x.shape[0])
        self._check(L.lib().mgpu_sg_set_data(self.ptr, _f(self._x), _f(self._y), x.shape[0]))
        self._x = self._y = None

    def train_step(self, idx: np.ndarray, lr: float, weight_decay: float = 0.0):
        idx = np.ascontiguousarray(idx, np.uint32)
        loss, ms = C.c_float(), C.c_float()
        self._check(L.lib().mgpu_sg_train_step(self.ptr, idx.ctypes.data_as(L.U32P), len(idx), lr,
                                               weight_decay, C.byref(loss), C.byref(ms)))
        return float(loss.value), float(ms.value)

    def predict_timed(self, n: int, repeats: int) -> float:
        ms = C.c_float()
        self._check(L.lib().mgpu_sg_predict_timed(self.ptr, n, repeats, C.byref(ms)))
        return float(ms.value)

    def close(self):
        if self.ptr:
            L.lib().mgpu_sg_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mfma_probe(ctx, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """One 32 x 32 x 16 bf16 MFMA through the kernels' lane layout; a, b are bf16 bit patterns."""
    a = np.ascontiguousarray(a, np.uint16)
    b = np.ascontiguousarray(b, np.uint16)
    d = np.zeros((32, 32), np.float32)
    rc = L.lib().mgpu_debug_mfma(ctx.ptr, a.ctypes.data_as(C.POINTER(C.c_uint16)),
                                 b.ctypes.data_as(C.POINTER(C.c_uint16)), _f(d))
    L.check(rc, ctx.ptr)
    return d


def to_bf16_bits(x: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(x, np.float32).view(np.uint32)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16_bits(h: np.ndarray) -> np.ndarray:
    return (h.astype(np.uint32) << 16).view(np.float32)


def bf16_round(x: np.ndarray) -> np.ndarray:
    return from_bf16_bits(to_bf16_bits(x))


# ---------------------------------------------------------------------------------------------
# data, training and evaluation
# ---------------------------------------------------------------------------------------------
def make_dataset(ctx, nchrom: int, seed: int, ncells: int = 512, contact_density: float = 30.0,
                 cell_seed: int | None = None, log=None):
    """Windows tiling nchrom simulated 20 Mbp chromosomes: (x, y = normalised log targets, raw
    counts, device seconds). Layouts come from `seed`, the cells' random streams from
    `cell_seed` (default: seed)."""
    from . import windows as Wn

    specs = Wn.sample_chromosomes(nchrom, seed)
    x, counts, means, dev_s = Wn.simulate_chromosomes(ctx, specs, ncells, contact_density,
                                                      seed if cell_seed is None else cell_seed,
                                                      log=log)
    return x, Wn.normalise(counts, means), counts, dev_s


def pair_correlations(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per-window mean per-diagonal correlation of two sets of count bands."""
    from . import sweep as Sw
    from . import windows as Wn

    m = Wn.valid_mask()
    return np.array([Sw.diag_correlation(np.where(m, u, 0), np.where(m, v, 0))
                     for u, v in zip(a, b)])


def train(net: Surrogate, x: np.ndarray, y: np.ndarray, epochs: float, batch: int, lr: float,
          weight_decay: float = 1e-4, warmup: int = 100, seed: int = 0, log=print):
    """AdamW with linear warm-up and cosine decay over epochs * n / batch steps of random
    batches; returns the per-step losses and the device milliseconds of all steps."""
    net.set_data(x, y)
    n = len(x)
    steps = max(1, int(epochs * n / batch))
    rng = np.random.default_rng(seed)
    losses, dev_ms = [], 0.0
    for t in range(steps):
        lr_t = lr * min(1.0, (t + 1) / warmup) * 0.5 * (1 + np.cos(np.pi * t / steps))
        idx = rng.choice(n, size=min(batch, n), replace=False).astype(np.uint32)
        loss, ms = net.train_step(idx, float(lr_t), weight_decay)
        losses.append(loss)
        dev_ms += ms
        if log and (t % max(1, steps // 20) == 0 or t == steps - 1):
            log(f"  step {t + 1}/{steps}: loss {np.mean(losses[-50:]):.4f}, lr {lr_t:.2e}")
    return np.array(losses), dev_ms


def window_correlations(net: Surrogate, x: np.ndarray, counts: np.ndarray) -> np.ndarray:
    """Mean per-diagonal correlation (25 kb, diagonals up to 500 kb) of each window's prediction
    with its simulated counts."""
    from . import sweep as Sw
    from . import windows as Wn

    return pair_correlations(Wn.denormalise(net.forward(x)), counts)


def save_params(net: Surrogate, path: str):
    np.save(path, net.params())


def load_params(net: Surrogate, path: str):
    p = np.load(path, allow_pickle=False)
    net.set_params(p)


def main(argv=None):
    import argparse
    import json
    import time

    from . import simulation as S

    ap = argparse.ArgumentParser(description="generate simulated windows and train the surrogate")
    ap.add_argument("--chromosomes", type=int, default=600,
                    help="simulated 20 Mbp chromosomes (13 windows each)")
    ap.add_argument("--val", type=int, default=30, help="validation chromosomes")
    ap.add_argument("--cells", type=int, default=512)
    ap.add_argument("--density", type=float, default=300.0, help="target contact density")
    ap.add_argument("--epochs", type=float, default=10.0)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None, help="weights (.npy); required for training")
    ap.add_argument("--summary", default=None, help="JSON summary")
    ap.add_argument("--time-forward", type=int, default=0, metavar="WINDOWS",
                    help="only time forward passes over a batch of WINDOWS windows with the "
                         "shipped weights (one untimed pass, then --repeats timed ones)")
    ap.add_argument("--repeats", type=int, default=20)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    ctx = S.Context(0)
    if a.time_forward:
        import os

        net = Surrogate(ctx, default_config(max_batch=a.time_forward))
        load_params(net, os.path.join(os.path.dirname(__file__), "data", "surrogate_weights.npy"))
        cfg = net.cfg
        x = np.random.default_rng(0).standard_normal(
            (a.time_forward, int(cfg.in_len), int(cfg.in_ch))).astype(np.float32)
        net.forward(x)
        ms = net.predict_timed(a.time_forward, a.repeats) / a.repeats
        log(f"forward of {a.time_forward} windows: {ms:.2f} ms")
        net.close()
        ctx.close()
        return
    if not a.out:
        ap.error("--out is required for training")
    t0 = time.perf_counter()
    log(f"simulating {a.chromosomes} training + {a.val} validation chromosomes, {a.cells} cells")
    x, y, _, dev_tr = make_dataset(ctx, a.chromosomes, a.seed, a.cells, a.density, log=log)
    xv, yv, cv, dev_va = make_dataset(ctx, a.val, a.seed + 1000, a.cells, a.density, log=log)
    t_data = time.perf_counter() - t0
    log(f"data: {len(x)} + {len(xv)} windows, {t_data:.1f} s wall, {dev_tr + dev_va:.1f} s in "
        f"the simulation kernel, {(a.chromosomes + a.val) / (dev_tr + dev_va):.1f} chromosomes/s")
    # noise ceiling: the validation layouts simulated again with other cell streams
    _, _, cv2, _ = make_dataset(ctx, a.val, a.seed + 1000, a.cells, a.density,
                                cell_seed=a.seed + 2000)
    ceil = pair_correlations(cv2, cv)
    nbar = xv[:, 100:500, 0].sum(axis=1)
    a.windows = len(x)
    log(f"noise ceiling (simulation vs simulation): mean {ceil.mean():.4f}; "
        f"windows with >= 20 barriers {ceil[nbar >= 20].mean():.4f}")
    net = Surrogate(ctx, default_config(max_batch=a.batch, seed=a.seed + 1))
    t0 = time.perf_counter()
    losses, dev_ms = train(net, x, y, a.epochs, a.batch, a.lr, seed=a.seed, log=log)
    t_train = time.perf_counter() - t0
    corr = window_correlations(net, xv, cv)
    log(f"validation: mean per-diagonal correlation {corr.mean():.4f} "
        f"(median {np.median(corr):.4f}, 5th percentile {np.percentile(corr, 5):.4f}); "
        f"windows with >= 20 barriers {corr[nbar >= 20].mean():.4f}")
    save_params(net, a.out)
    if a.summary:
        with open(a.summary, "w") as fh:
            json.dump({"windows": a.windows, "cells": a.cells, "contact_density": a.density,
                       "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "params": net.nparams,
                       "data_wall_s": t_data, "data_kernel_s": dev_tr + dev_va,
                       "train_wall_s": t_train, "train_device_s": dev_ms / 1000.0,
                       "final_loss": float(np.mean(losses[-50:])),
                       "val_corr_mean": float(corr.mean()), "val_corr_median": float(np.median(corr)),
                       "val_corr_p5": float(np.percentile(corr, 5)),
                       "val_corr_20bar": float(corr[nbar >= 20].mean()),
                       "ceiling_mean": float(ceil.mean()),
                       "ceiling_20bar": float(ceil[nbar >= 20].mean())}, fh, indent=1)
    net.close()
    ctx.close()


if __name__ == "__main__":
    main()
