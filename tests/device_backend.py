"""KAT-replay backend over the device step code (mgpu_debug_step through the C-ABI)."""
import ctypes as C

import numpy as np

from modlefork_amd import _lib as L
from modlefork_amd import simulation as S

OPS = {1: 1, 3: 3, 4: 4, 5: 5, 6: 6, 7: 7, 8: 8}


def _p(a, ct=C.c_uint64):
    return a.ctypes.data_as(C.POINTER(ct))


def params_from_rc(cfg: dict) -> L.MgpuParams:
    p = L.MgpuParams()
    names = {n for n, _ in L.MgpuParams._fields_}
    for k, v in cfg.items():
        if k in names:
            setattr(p, k, v)
    return p


class DeviceBackend:
    def __init__(self, ctx: S.Context):
        self.ctx = ctx

    def begin(self, cfg, chrom, st):
        self.params = params_from_rc(cfg)
        self.chrom = S.chrom_view(chrom["name"], chrom["size"], chrom["start"], chrom["end"])
        self.bars = S.Barriers(st["bar_pos"], st["bar_dir"], st["bar_stp_active"],
                               st["bar_stp_inactive"])
        self.bst = st["bar_state"] if len(st["bar_state"]) else np.zeros(1, np.uint8)
        from oracle import refcpu as R  # engine seeding is checked separately; reuse the state
        g = R.rng(st["seed"])
        self.rng = np.array(list(g.s), np.uint64)

    def _step(self, st, op):
        n = len(st["rev"])
        self.ctx.check(L.lib().mgpu_debug_step(
            self.ctx.ptr, C.byref(self.params), C.byref(self.chrom), op, n, _p(st["rev"]),
            _p(st["fwd"]), _p(st["epoch"]), _p(st["rev_ranks"]), _p(st["fwd_ranks"]),
            _p(st["rev_moves"]), _p(st["fwd_moves"]), _p(st["rev_coll"]), _p(st["fwd_coll"]),
            C.byref(self.bars.view), _p(self.bst, C.c_uint8), _p(self.rng)))

    def adjust_and_clamp(self, st):
        self._step(st, 0)

    def collisions(self, st, mode):
        self._step(st, OPS[mode])

    def rank(self, rev, fwd, ep, rr, fr, partially_sorted, init):
        st = dict(rev=rev, fwd=fwd, epoch=ep, rev_ranks=rr, fwd_ranks=fr,
                  rev_moves=np.zeros(len(rev), np.uint64), fwd_moves=np.zeros(len(rev), np.uint64),
                  rev_coll=np.zeros(len(rev), np.uint64), fwd_coll=np.zeros(len(rev), np.uint64))
        if not init:
            raise NotImplementedError
        self._step(st, 10 if partially_sorted else 9)
