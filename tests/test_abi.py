"""The C-ABI library: it loads, exports every entry point of include/modle_gpu.h, and its
host-only functions (parameter transforms, barrier preparation, planning, per-cell seeds) agree
with the oracle's restatement of the reference. No GPU is needed."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest
import xxhash

from modlefork_amd import _lib as L
from modlefork_amd import simulation as S
from modlefork_amd import workloads as W
from oracle import config as OC
from oracle import refcpu as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "modle_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mgpu_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    declared = header_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(n for n, _, _ in L.SIGNATURES) == declared


def test_cool_library_exports_every_declared_symbol():
    """libmodle_cool.so (the .cool writer, where HDF5 is available) exports every entry point of
    include/modle_cool.h."""
    import ctypes

    from modlefork_amd import cooler as CL

    if not os.path.exists(CL.LIB_PATH):
        pytest.skip("HDF5 is not available: the .cool writer is not built")
    text = open(os.path.join(ROOT, "include", "modle_cool.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    declared = sorted(set(re.findall(r"\b(mcool_\w+)\s*\(", text)))
    assert declared == ["mcool_close", "mcool_last_error", "mcool_open", "mcool_write_chrom",
                        "mcool_write_metadata"]
    lib = ctypes.CDLL(CL.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name


def test_version_string():
    assert b"modle" in L.lib().mgpu_version()


OPTION_SETS = [
    W.small().options,
    W.genome().options,
    W.stress().options,
    dict(bin_size=5000, sample_tad_contacts=0),
    dict(bin_size=2000, rev_extrusion_speed=1000, probability_of_extrusion_unit_bypass=0.3),
    dict(bin_size=10000, lef_bar_minor_collision_pblock=0.5, stopping_criterion=1),
]

ORACLE_NAMES = {"sample_tad_contacts": "sample_tad", "sample_loop_contacts": "sample_loop",
                "noisify_contacts": "noisify"}


def oracle_transform(opts):
    o = {}
    for k, v in opts.items():
        if k == "stopping_criterion":
            o["stop_on_epochs"] = v == 1
        else:
            o[ORACLE_NAMES.get(k, k)] = bool(v) if k in ORACLE_NAMES else v
    return OC.transform(o)


@pytest.mark.parametrize("opts", OPTION_SETS)
def test_config_transform_matches_oracle(opts):
    p = S.make_params({k: v for k, v in opts.items() if k != "num_cells"}).as_dict()
    c = oracle_transform(opts)
    rc = OC.to_rc(c)
    for k, v in rc.items():
        if isinstance(v, float) and math.isinf(v):
            assert math.isinf(p[k]), k
        else:
            assert p[k] == v, (k, p[k], v)
    assert p["barrier_occupied_stp"] == c["barrier_occupied_stp"]
    assert p["barrier_not_occupied_stp"] == c["barrier_not_occupied_stp"]


def test_default_parameters_closed_forms():
    # 5 kb bins: v = 0.8 * bin, sigma = 0.05 v, p_release = (v_r + v_f) / processivity,
    # pbb from occupancy 0.825 with puu 0.7, activation epochs = 5 * processivity / (v_r + v_f)
    p = S.make_params(W.small().options).as_dict()
    assert p["rev_speed"] == p["fwd_speed"] == 4000
    assert p["rev_speed_std"] == 200.0
    assert p["prob_release"] == 8000 / 300000
    assert abs(p["barrier_occupied_stp"] - (1 - 0.3 * (1 - 0.825) / 0.825)) < 1e-15
    assert p["burnin_target_epochs_for_lef_activation"] == 187
    # 1 kb bins normalise by (800 + 800) / 8000 = 0.2
    q = S.make_params(W.stress().options).as_dict()
    assert abs(q["p_bypass"] - 0.02) < 1e-15
    assert abs(q["barrier_not_occupied_stp"] - 0.7 ** 0.2) < 1e-12


@pytest.mark.parametrize("wl", [W.small(), W.stress(ncells=4)])
def test_prepare_barriers_matches_oracle(wl):
    name, size = wl.chroms[0]
    p = S.make_params({k: v for k, v in wl.options.items() if k != "num_cells"})
    b = S.prepare_barriers(p, wl.barriers_of(name), 0, size)
    o = OC.prepare_barriers(wl.barriers_of(name), 0, size, OC.transform(wl.options))
    assert np.array_equal(b.pos, o.pos)
    assert np.array_equal(b.dir, o.dir)
    assert np.array_equal(b.stp_active, o.stp_active)
    assert np.array_equal(b.stp_inactive, o.stp_inactive)
    assert np.all(np.diff(b.pos.astype(np.int64)) >= 0)


def test_barriers_at_equal_midpoints_follow_pdq_sort():
    """Barriers whose BED records share a midpoint: ExtrusionBarriers::sort orders the indices
    with cpp-sort's pdq_sort, which is unstable (extrusion_barriers.cpp:239-259). The library's
    order equals the oracle's pdq restatement field by field, on a set where that order differs
    from the records' own (a stable sort's) order, and on one small enough for pdq_sort's
    insertion-sort path (n < 24), where the two agree."""
    rng = np.random.default_rng(12)
    for n in (300, 20):
        mids = rng.integers(0, 12, n) * 1000 + 500
        recs = [W.BedRecord("c", int(m) - 1, int(m), float(rng.uniform(0.05, 0.95)),
                            "+" if rng.random() < 0.5 else "-") for m in mids]
        opts = dict(bin_size=1000, diagonal_width=10_000)
        p = S.make_params(opts)
        b = S.prepare_barriers(p, recs, 0, 20_000)
        o = OC.prepare_barriers(recs, 0, 20_000, OC.transform(opts))
        assert np.array_equal(b.pos, o.pos) and np.array_equal(b.dir, o.dir)
        assert np.array_equal(b.stp_active, o.stp_active)
        assert np.array_equal(b.stp_inactive, o.stp_inactive)
        assert np.all(np.diff(b.pos.astype(np.int64)) >= 0)
        stable = sorted(range(n), key=lambda i: (recs[i].start + recs[i].end + 1) // 2)
        stable_sa = np.array([OC.barrier_stp(recs[i].score, OC.transform(opts))[0] for i in stable])
        if n >= 24:
            assert not np.array_equal(b.stp_active, stable_sa)
        else:
            assert np.array_equal(b.stp_active, stable_sa)


def test_prepare_barriers_rejects_bad_records():
    p = S.make_params({})
    bad = [W.BedRecord("c", 10, 11, 1.5, "+")]
    with pytest.raises(L.MgpuError):
        S.prepare_barriers(p, bad, 0, 100)
    bad = [W.BedRecord("c", 10, 11, 0.5, "x")]
    with pytest.raises(L.MgpuError):
        S.prepare_barriers(p, bad, 0, 100)


def test_bed_name_field_as_not_bound_stp():
    """--interpret-extrusion-barrier-name-as-not-bound-stp (genome.cpp:459-468): each record's
    name field is its not-bound stp. With a score the active stp is the one whose stationary
    occupancy is that score (the closed form of extrusion_barriers_impl.hpp:106-128, checked
    here by inverting it); without one it is the configured default. The occupancy override
    (simulation.cpp:68-78) restores the configured pair. A name outside [0, 1] is an error even
    for a record outside the simulated range, as the reference checks before dropping it."""
    rng = np.random.default_rng(5)
    n = 200
    recs = [W.BedRecord("c", int(m), int(m) + 1, float(rng.choice([0.0, rng.uniform(0.05, 0.95)])),
                        "+" if rng.random() < 0.5 else "-")
            for m in rng.choice(100_000, n, replace=False)]
    puu = rng.uniform(0.05, 0.95, n)
    opts = dict(bin_size=1000, diagonal_width=10_000)
    c = OC.transform(opts)
    p = S.make_params(opts)
    b = S.prepare_barriers(p, recs, 0, 100_000, name_puu=puu)
    o = OC.prepare_barriers(recs, 0, 100_000, c, name_puu=puu)
    for f in ("pos", "dir", "stp_active", "stp_inactive"):
        assert np.array_equal(getattr(b, f), getattr(o, f)), f
    by_pos = {r.start + 1: (r, u) for r, u in zip(recs, puu)}  # midpoint of a 1-bp record
    for pos, sa, si in zip(b.pos, b.stp_active, b.stp_inactive):
        r, u = by_pos[int(pos)]
        assert si == u
        if r.score != 0.0:
            # the closed form clamps stp_active to [0, 1]: an occupancy below the one an
            # always-leaving active state gives, (1 - u) / (2 - u), maps to 0
            if sa > 0.0:
                assert abs(L.lib().mgpu_occupancy_from_stp(sa, si) - r.score) < 1e-12
            else:
                assert r.score <= (1 - u) / (2 - u) + 1e-12
        else:
            assert sa == p.barrier_occupied_stp
    # without the name field the default not-bound stp applies
    d = S.prepare_barriers(p, recs, 0, 100_000)
    assert np.all(d.stp_inactive == p.barrier_not_occupied_stp)
    # the override restores the configured pair
    po = S.make_params(dict(opts, extrusion_barrier_occupancy=0.9))
    assert po.override_barrier_occupancy == 1
    ov = S.prepare_barriers(po, recs, 0, 100_000, name_puu=puu)
    assert np.all(ov.stp_active == po.barrier_occupied_stp)
    assert np.all(ov.stp_inactive == po.barrier_not_occupied_stp)
    # an out-of-range name fails, also on a record outside the chromosome range
    bad = [W.BedRecord("c", 10, 11, 0.5, "+"), W.BedRecord("c", 500, 501, 0.5, "+")]
    with pytest.raises(L.MgpuError):
        S.prepare_barriers(p, bad, 0, 100, name_puu=[0.5, 1.5])
    with pytest.raises(ValueError):
        OC.prepare_barriers(bad, 0, 100, c, name_puu=[0.5, 1.5])
    assert len(S.prepare_barriers(p, bad, 0, 100, name_puu=[0.5, 0.25]).pos) == 1


def test_prepare_barriers_drops_out_of_range_and_uses_midpoint():
    p = S.make_params({})
    recs = [W.BedRecord("c", 90, 100, 0.0, "+"), W.BedRecord("c", 10, 13, 0.0, "-"),
            W.BedRecord("c", 4, 5, 0.0, "+")]
    b = S.prepare_barriers(p, recs, 5, 95)
    # midpoints (start + end + 1) / 2 = 95 (outside [5, 95), dropped), 12 and 5; sorted by position
    assert b.pos.tolist() == [5, 12]
    assert b.dir.tolist() == [1, 2]


@pytest.mark.parametrize("wl", [W.small(), W.genome(), W.stress()])
def test_plan_matches_oracle(wl):
    c = OC.transform(wl.options)
    rc = R.make_config(OC.to_rc(c))
    p = S.make_params({k: v for k, v in wl.options.items() if k != "num_cells"})
    ncells = wl.options["num_cells"]
    for name, size in wl.chroms[:4]:
        ch = S.chrom_view(name, size)
        plan = S.plan_chrom(p, ch, ncells)
        nl = R.lib().rc_num_lefs(c["number_of_lefs_per_mbp"], size)
        ncols = R.lib().rc_band_ncols(c["bin_size"], size)
        nrows = R.lib().rc_band_nrows(c["diagonal_width"], c["bin_size"], size)
        t = np.zeros(ncells, np.uint64)
        te = R.lib().rc_plan_targets(C.byref(rc), ncells, nl, nrows * ncols,
                                     t.ctypes.data_as(C.POINTER(C.c_uint64)))
        assert (plan.nlefs, plan.nrows, plan.ncols, plan.target_epochs) == (nl, nrows, ncols, te)
        assert np.array_equal(plan.targets, t)
        assert int(plan.targets.sum()) == round(nrows * ncols * c["target_contact_density"])


def test_cell_seed_matches_xxh3_and_oracle():
    for name, size in [("chrS", 20_000_000), ("chr1", 248_956_422), ("x" * 100, 7), ("", 1)]:
        ch = S.chrom_view(name, size)
        for cell in (0, 1, 63, 511, 2**40):
            for seed in (0, 1, 2**63 + 5):
                data = name.encode() + size.to_bytes(8, "little") + cell.to_bytes(8, "little")
                want = xxhash.xxh3_64_intdigest(data, seed=seed)
                assert L.lib().mgpu_cell_seed(ch, cell, seed) == want
                assert R.lib().rc_cell_seed(name.encode(), len(name), size, cell, seed) == want


def test_no_device_is_reported_as_error_not_fallback():
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a device is present")
    with pytest.raises(L.MgpuError):
        S.Context(0)


def _abi_caller():
    import os
    import subprocess

    here = os.path.dirname(os.path.abspath(__file__))
    exe = os.path.join(here, "cpp", "abi_caller")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(here, "cpp")], check=True)
    return exe


def test_cpp_caller_fills_params_like_the_transform():
    """A C++ program compiled against include/modle_gpu.h and include/modle_gpu_config.hpp (no
    ctypes): fill_params_from_config on a modle::Config holding what Cli::transform_args leaves
    for `-r 5kb -w 1Mbp --extrusion-barrier-occupancy 0.825` gives mgpu_config_transform's block
    for the same options, and both hold the defaults' closed forms (4 kb per epoch, sigma 200 bp,
    release 8000 / 300000, 187 activation epochs, occupied stp 1 - 0.3 * 0.175 / 0.825)."""
    import subprocess

    r = subprocess.run([_abi_caller(), "params"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr
