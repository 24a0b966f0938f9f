"""Replays the reference's known-answer unit tests (tests/golden/reference_kat.json) through a
backend: the CPU oracle, or the device step kernels behind the C-ABI.

A backend provides
    adjust_and_clamp(cfg, chrom, state)
    collisions(cfg, chrom, state, mode)     # modes as in oracle/refcpu.h rc_process_collisions
    rank(state, lefs, partially_sorted, init)
where `state` is a dict of numpy arrays (rev, fwd, epoch, rev_ranks, fwd_ranks, rev_moves,
fwd_moves, rev_coll, fwd_coll) plus barrier arrays and the engine seed.
"""
from __future__ import annotations

import json
import os

import numpy as np

UNBOUND = np.uint64((1 << 64) - 1)
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_kat.json")

HOOK_MODE = {
    "test_process_collisions": 1,
    "test_fix_secondary_lef_lef_collisions": 4,
    "test_detect_lef_bar_collisions": 5,
    "test_detect_primary_lef_lef_collisions": 6,
    "test_process_lef_lef_collisions": 3,
    "test_correct_moves_for_lef_bar_collisions": 7,
    "test_detect_units_at_chrom_boundaries": 8,
}

BASE_CONFIG = dict(
    bin_size=5000, diagonal_width=3_000_000, rev_speed=4000, fwd_speed=4000,
    rev_speed_burnin=4000, fwd_speed_burnin=4000, rev_speed_std=0.0, fwd_speed_std=0.0,
    prob_release=0.0, prob_release_burnin=0.0, p_bypass=0.1, pblock_major=1.0,
    pblock_minor=0.0, hard_stall_mult=5.0, soft_stall_mult=1.0, tad_to_loop_ratio=5.0,
    gev_mu=0.0, gev_sigma=5000.0, gev_xi=0.001, noisify=1, skip_burnin=0,
    contact_sampling_interval=50_000, burnin_history_length=100, burnin_smoothing_window=5,
    min_burnin_epochs=0, max_burnin_epochs=(1 << 64) - 1,
    burnin_target_epochs_for_lef_activation=320, target_contact_density=1.0,
    target_simulation_epochs=2000, seed=0,
)


def load_cases():
    with open(GOLDEN) as fh:
        return json.load(fh)


def kat_cases():
    """Cases with exact expected values (the property-only tests are excluded)."""
    out = []
    for c in load_cases():
        arrs = c["arrays"]
        if any(k.endswith("_expected") or k.endswith("_adjusted") or k.startswith("rev_ranks_expected")
               for k in arrs) and c["hooks"]:
            out.append(c)
    return out


def encode(pair):
    idx, ev = pair
    return np.uint64(idx | (ev << 56)) if ev else np.uint64(0)


def lef_arrays(values):
    n = len(values)
    rev = np.zeros(n, np.uint64)
    fwd = np.zeros(n, np.uint64)
    ep = np.zeros(n, np.uint64)
    for i, v in enumerate(values):
        if v is None:
            rev[i] = fwd[i] = ep[i] = UNBOUND
        else:
            rev[i], fwd[i], ep[i] = v
    return rev, fwd, ep


def make_state(case):
    a = {k: v["values"] for k, v in case["arrays"].items()}
    st = {}
    if "lefs" in a:
        st["rev"], st["fwd"], st["epoch"] = lef_arrays(a["lefs"])
        n = len(a["lefs"])
        st["rev_ranks"] = np.array(a.get("rev_ranks", list(range(n))), np.uint64)
        st["fwd_ranks"] = np.array(a.get("fwd_ranks", list(range(n))), np.uint64)
        st["rev_moves"] = np.array(a.get("rev_moves", [0] * n), np.uint64)
        st["fwd_moves"] = np.array(a.get("fwd_moves", [0] * n), np.uint64)
        st["rev_coll"] = np.zeros(n, np.uint64)
        st["fwd_coll"] = np.zeros(n, np.uint64)
    bars = case.get("barriers") or []
    st["bar_pos"] = np.array([b["pos"] for b in bars], np.uint64)
    st["bar_dir"] = np.array([b["dir"] for b in bars], np.uint8)
    st["bar_stp_active"] = np.array([b["stp_active"] for b in bars], np.float64)
    st["bar_stp_inactive"] = np.array([b["stp_inactive"] for b in bars], np.float64)
    st["bar_state"] = np.array([b["state"] for b in bars], np.uint8)
    st["seed"] = case["seed"]
    return st


def case_config(case):
    cfg = dict(BASE_CONFIG)
    for k, v in case["config"].items():
        if k in cfg and not isinstance(v, str):
            cfg[k] = type(cfg[k])(v)
    return cfg


def case_chrom(case):
    ch = case.get("chrom") or dict(name="chr1", start=0, end=1 << 62, size=1 << 62)
    return ch


def replay(case, backend):
    """Runs the case's hooks in order and returns a list of (what, got, expected) mismatches."""
    cfg = case_config(case)
    chrom = case_chrom(case)
    st = make_state(case)
    a = {k: v["values"] for k, v in case["arrays"].items()}
    bad = []
    backend.begin(cfg, chrom, st)
    for h in case["hooks"]:
        name = h["hook"]
        if name == "test_adjust_and_clamp_moves":
            backend.adjust_and_clamp(st)
        elif name in HOOK_MODE:
            backend.collisions(st, HOOK_MODE[name])
        elif name == "test_rank_lefs":
            lefs_name = h["args"][0]
            rev, fwd, ep = lef_arrays(a[lefs_name])
            n = len(rev)
            rr = np.zeros(n, np.uint64)
            fr = np.zeros(n, np.uint64)
            backend.rank(rev, fwd, ep, rr, fr, partially_sorted=False, init=True)
            suffix = lefs_name[-1]
            for which, got in (("rev", rr), ("fwd", fr)):
                exp = np.array(a[f"{which}_ranks_expected{suffix}"], np.uint64)
                if not np.array_equal(got, exp):
                    bad.append((f"{which}_ranks_{suffix}", got.tolist(), exp.tolist()))
        else:
            raise ValueError(name)
    for which in ("rev", "fwd"):
        key = f"{which}_collisions_expected"
        if key in a:
            exp = np.array([encode(p) for p in a[key]], np.uint64)
            if not np.array_equal(st[f"{which}_coll"], exp):
                bad.append((key, st[f"{which}_coll"].tolist(), exp.tolist()))
        for key in (f"{which}_moves_expected", f"{which}_moves_adjusted"):
            if key in a:
                exp = np.array(a[key], np.uint64)
                if not np.array_equal(st[f"{which}_moves"], exp):
                    bad.append((key, st[f"{which}_moves"].tolist(), exp.tolist()))
    for arr, i, v in case["rank_checks"]:
        if int(st[arr][i]) != v:
            bad.append((f"{arr}[{i}]", int(st[arr][i]), v))
    return bad
