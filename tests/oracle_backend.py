"""KAT-replay backend over the CPU oracle (test infrastructure)."""
import ctypes as C

import numpy as np

from oracle import refcpu as R


def _p(a, ct=C.c_uint64):
    return a.ctypes.data_as(C.POINTER(ct))


class OracleBackend:
    def begin(self, cfg, chrom, st):
        self.cfg = R.make_config(cfg)
        self.chrom = R.make_chrom(chrom["name"], chrom["size"], chrom["start"], chrom["end"])
        self.bars = R.Barriers(st["bar_pos"], st["bar_dir"], st["bar_stp_active"],
                               st["bar_stp_inactive"])
        self.g = R.rng(st["seed"])

    def adjust_and_clamp(self, st):
        n = len(st["rev"])
        R.lib().rc_adjust_and_clamp_moves(C.byref(self.chrom), _p(st["rev"]), _p(st["fwd"]),
                                          _p(st["epoch"]), n, _p(st["rev_ranks"], C.c_size_t),
                                          _p(st["fwd_ranks"], C.c_size_t), _p(st["rev_moves"]),
                                          _p(st["fwd_moves"]), 1)

    def collisions(self, st, mode):
        n = len(st["rev"])
        bst = st["bar_state"] if len(st["bar_state"]) else np.zeros(1, np.uint8)
        R.lib().rc_process_collisions(C.byref(self.cfg), C.byref(self.chrom), _p(st["rev"]),
                                      _p(st["fwd"]), _p(st["epoch"]), n,
                                      _p(st["rev_ranks"], C.c_size_t),
                                      _p(st["fwd_ranks"], C.c_size_t), _p(st["rev_moves"]),
                                      _p(st["fwd_moves"]), C.byref(self.bars.c),
                                      _p(bst, C.c_uint8), _p(st["rev_coll"]), _p(st["fwd_coll"]),
                                      C.byref(self.g), mode, None, None)

    def rank(self, rev, fwd, ep, rr, fr, partially_sorted, init):
        R.lib().rc_rank_lefs(_p(rev), _p(fwd), _p(ep), len(rev), _p(rr, C.c_size_t),
                             _p(fr, C.c_size_t), int(partially_sorted), int(init))
