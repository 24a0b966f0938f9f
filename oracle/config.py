"""Oracle-side restatement of MoDLE's simulate parameters and their CLI-derived values.

TEST INFRASTRUCTURE: the checker, never the product.

Defaults follow src/common/include/modle/common/simulation_config.hpp:24-135; the derived values
follow Cli::transform_args, src/modle/cli.cpp:1085-1215. `parsed` names the options a user passed
explicitly (the CLI's "was this option given" tests).
"""
from __future__ import annotations

import math

from . import refcpu as R

DEFAULTS = dict(
    bin_size=5_000,
    diagonal_width=3_000_000,
    tad_to_loop_contact_ratio=5.0,
    genextreme_mu=0.0,
    genextreme_sigma=5_000.0,
    genextreme_xi=0.001,
    sample_loop=True,
    sample_tad=True,
    noisify=True,
    rev_extrusion_speed=None,  # 0.8 * bin size unless given
    fwd_extrusion_speed=None,
    rev_extrusion_speed_std=0.05,
    fwd_extrusion_speed_std=0.05,
    number_of_lefs_per_mbp=20.0,
    avg_lef_processivity=300_000,
    contact_sampling_interval=50_000,
    extrusion_barrier_occupancy=0.825,
    barrier_occupied_stp=0.0,
    barrier_not_occupied_stp=0.70,
    hard_stall_lef_stability_multiplier=5.0,
    soft_stall_lef_stability_multiplier=1.0,
    probability_of_extrusion_unit_bypass=0.1,
    lef_bar_major_collision_pblock=1.0,
    lef_bar_minor_collision_pblock=0.0,
    target_contact_density=1.0,
    target_simulation_epochs=2000,
    stop_on_epochs=False,
    num_cells=512,
    seed=0,
    probability_normalization_factor=8_000,
    normalize_probabilities=True,
    skip_burnin=False,
    burnin_history_length=100,
    burnin_smoothing_window_size=5,
    min_burnin_epochs=0,
    max_burnin_epochs=(1 << 64) - 1,
    burnin_speed_coefficient=1.0,
)


def stp_active_from_occupancy(stp_inactive: float, occupancy: float) -> float:
    return R.lib().rc_stp_active_from_occupancy(stp_inactive, occupancy)


def occupancy_from_stp(stp_active: float, stp_inactive: float) -> float:
    return R.lib().rc_occupancy_from_stp(stp_active, stp_inactive)


def transform(user: dict | None = None) -> dict:
    """Returns the full parameter dict after the CLI transforms (cli.cpp:1192-1215)."""
    c = dict(DEFAULTS)
    user = dict(user or {})
    # the C-ABI's field names (modle::Config's) for the three switches
    for field, name in (("noisify_contacts", "noisify"), ("sample_loop_contacts", "sample_loop"),
                        ("sample_tad_contacts", "sample_tad")):
        if field in user:
            user[name] = bool(user.pop(field))
    parsed = set(user)
    c.update(user)
    # cli_update_extr_speed (cli.cpp:1085-1110)
    if "rev_extrusion_speed" not in parsed or c["rev_extrusion_speed"] is None:
        c["rev_extrusion_speed"] = c["bin_size"] * 8 // 10
    if "fwd_extrusion_speed" not in parsed or c["fwd_extrusion_speed"] is None:
        c["fwd_extrusion_speed"] = c["bin_size"] * 8 // 10
    if 0 < c["fwd_extrusion_speed_std"] < 1:
        c["fwd_extrusion_speed_std"] *= float(c["fwd_extrusion_speed"])
    if 0 < c["rev_extrusion_speed_std"] < 1:
        c["rev_extrusion_speed_std"] *= float(c["rev_extrusion_speed"])
    c["rev_extrusion_speed_burnin"] = int(_round(c["burnin_speed_coefficient"] * float(c["rev_extrusion_speed"])))
    c["fwd_extrusion_speed_burnin"] = int(_round(c["burnin_speed_coefficient"] * float(c["fwd_extrusion_speed"])))
    # cli_compute_prob_of_lef_release (:1112-1119)
    c["prob_of_lef_release"] = float(c["rev_extrusion_speed"] + c["fwd_extrusion_speed"]) / float(c["avg_lef_processivity"])
    c["prob_of_lef_release_burnin"] = float(c["rev_extrusion_speed_burnin"] + c["fwd_extrusion_speed_burnin"]) / float(c["avg_lef_processivity"])
    # cli_update_barrier_stp_and_occupancy (:1122-1135)
    if "extrusion_barrier_occupancy" in parsed:
        c["barrier_occupied_stp"] = stp_active_from_occupancy(c["barrier_not_occupied_stp"], c["extrusion_barrier_occupancy"])
    else:
        c["extrusion_barrier_occupancy"] = occupancy_from_stp(c["barrier_occupied_stp"], c["barrier_not_occupied_stp"])
    # cli_update_tad_to_loop_contact_ratio (:1169-1182)
    if c["sample_loop"] and not c["sample_tad"]:
        c["tad_to_loop_contact_ratio"] = 0.0
    if not c["sample_loop"] and c["sample_tad"]:
        c["tad_to_loop_contact_ratio"] = math.inf
    # cli_update_burnin_params (:1184-1190)
    act_bp = 5 * c["avg_lef_processivity"]
    c["burnin_target_epochs_for_lef_activation"] = min(
        c["max_burnin_epochs"], act_bp // (c["rev_extrusion_speed_burnin"] + c["fwd_extrusion_speed_burnin"]))
    # cli_normalize_probabilities (:1138-1167)
    if c["normalize_probabilities"]:
        ratio = float(c["rev_extrusion_speed"] + c["fwd_extrusion_speed"]) / float(c["probability_normalization_factor"])
        if ratio != 1.0:
            def stable_pow(base, e):
                if base == 0.0:
                    return 0.0
                if base == 1.0:
                    return 1.0
                return math.exp(math.log(base) * e)
            c["barrier_not_occupied_stp"] = stable_pow(c["barrier_not_occupied_stp"], ratio)
            c["barrier_occupied_stp"] = stp_active_from_occupancy(c["barrier_not_occupied_stp"], c["extrusion_barrier_occupancy"])
            p = c["probability_of_extrusion_unit_bypass"]
            if p != 0.0 and p != 1.0:
                c["probability_of_extrusion_unit_bypass"] = min(p * ratio, 1.0)
            c["lef_bar_major_collision_pblock"] = stable_pow(c["lef_bar_major_collision_pblock"], ratio)
            c["lef_bar_minor_collision_pblock"] = stable_pow(c["lef_bar_minor_collision_pblock"], ratio)
    c["override_extrusion_barrier_occupancy"] = "extrusion_barrier_occupancy" in parsed
    if c["stop_on_epochs"]:
        c["target_contact_density"] = -1.0
    return c


def _round(x: float) -> float:
    # std::round: half away from zero
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def to_rc(c: dict) -> dict:
    """Maps a transformed parameter dict onto the oracle's rc_config fields."""
    return dict(
        bin_size=c["bin_size"],
        diagonal_width=c["diagonal_width"],
        rev_speed=c["rev_extrusion_speed"],
        fwd_speed=c["fwd_extrusion_speed"],
        rev_speed_burnin=c["rev_extrusion_speed_burnin"],
        fwd_speed_burnin=c["fwd_extrusion_speed_burnin"],
        rev_speed_std=c["rev_extrusion_speed_std"],
        fwd_speed_std=c["fwd_extrusion_speed_std"],
        prob_release=c["prob_of_lef_release"],
        prob_release_burnin=c["prob_of_lef_release_burnin"],
        p_bypass=c["probability_of_extrusion_unit_bypass"],
        pblock_major=c["lef_bar_major_collision_pblock"],
        pblock_minor=c["lef_bar_minor_collision_pblock"],
        hard_stall_mult=c["hard_stall_lef_stability_multiplier"],
        soft_stall_mult=c["soft_stall_lef_stability_multiplier"],
        tad_to_loop_ratio=c["tad_to_loop_contact_ratio"],
        gev_mu=c["genextreme_mu"],
        gev_sigma=c["genextreme_sigma"],
        gev_xi=c["genextreme_xi"],
        noisify=1 if c["noisify"] else 0,
        skip_burnin=1 if c["skip_burnin"] else 0,
        contact_sampling_interval=c["contact_sampling_interval"],
        burnin_history_length=c["burnin_history_length"],
        burnin_smoothing_window=c["burnin_smoothing_window_size"],
        min_burnin_epochs=c["min_burnin_epochs"],
        max_burnin_epochs=c["max_burnin_epochs"],
        burnin_target_epochs_for_lef_activation=c["burnin_target_epochs_for_lef_activation"],
        target_contact_density=c["target_contact_density"],
        target_simulation_epochs=c["target_simulation_epochs"],
        seed=c["seed"],
    )


def barrier_stp(score: float, c: dict, puu=None):
    """compute_barrier_stp (genome.cpp:181-192) with the default not-bound stp, or `puu` when the
    BED name field gives it (genome.cpp:459-468), followed by the occupancy override of
    Simulation::Simulation (simulation.cpp:68-78). Returns (stp_active, stp_inactive)."""
    if c.get("override_extrusion_barrier_occupancy"):
        return c["barrier_occupied_stp"], c["barrier_not_occupied_stp"]
    si = c["barrier_not_occupied_stp"] if puu is None else puu
    if score != 0.0:
        return stp_active_from_occupancy(si, score), si
    return c["barrier_occupied_stp"], si


def import_barriers(records, chrom_start: int, chrom_end: int, c: dict, name_puu=None):
    """Genome::import_barriers (genome.cpp:433-480) for one chromosome's BED records in file
    order: the checks of every record, then the records ordered by start (the parser's interval
    tree, IITree::make_BST's pdq_sort, interval_tree_impl.hpp:58-68), those whose midpoint lies in
    [chrom_start, chrom_end) kept (Chromosome::add_extrusion_barrier, genome.cpp:194-210) and
    ordered by start again (the chromosome's own tree). Returns rows (bed_start, bed_end, pos,
    dir, stp_active, stp_inactive) in that order. name_puu, when given, holds each record's name
    field as a number (interpret_name_field_as_puu)."""
    for i, r in enumerate(records):
        if r.score < 0 or r.score > 1:
            raise ValueError(f"record {i}: score must be in [0, 1]")
        if name_puu is not None and not 0.0 <= float(name_puu[i]) <= 1.0:
            raise ValueError(f"record {i}: name field must be in [0, 1]")
    kept = [i for i in R.pdq_order([r.start for r in records])
            if chrom_start <= (records[i].start + records[i].end + 1) // 2 < chrom_end]
    rows = []
    for k in R.pdq_order([records[i].start for i in kept]):
        i = kept[k]
        r = records[i]
        sa, si = barrier_stp(r.score, c, None if name_puu is None else float(name_puu[i]))
        rows.append((r.start, r.end, (r.start + r.end + 1) // 2, 1 if r.strand == "+" else 2,
                     sa, si))
    return rows


def sort_barriers(rows) -> "R.Barriers":
    """The ExtrusionBarriers constructor's position sort (extrusion_barriers.cpp:234-259): cpp-sort
    pdq_sort of the indices by position; equal positions keep pdq_sort's (unstable) order.
    rows as import_barriers returns them."""
    rows = [rows[i] for i in R.pdq_order([t[2] for t in rows])]
    return R.Barriers([t[2] for t in rows], [t[3] for t in rows], [t[4] for t in rows],
                      [t[5] for t in rows])


def prepare_barriers(records, chrom_start: int, chrom_end: int, c: dict,
                     name_puu=None) -> "R.Barriers":
    """A chromosome's barrier tables from its BED records: import_barriers, then sort_barriers."""
    return sort_barriers(import_barriers(records, chrom_start, chrom_end, c, name_puu))