"""Seeded synthetic inputs for the simulate path: chrom.sizes entries plus BED6 barrier records.

Each workload is the input a user would hand to `modle simulate` (chromosomes, barrier BED
records and the CLI options that differ from the defaults); nothing here runs the simulation.

  * small:   one 20 Mbp chromosome, 200 uniform barriers, 5 kb bins, 3 Mbp band, 64 cells.
  * genome:  24 chromosomes of log-uniform length in [40, 250] Mbp, clustered barriers with
             Beta(8,2) occupancy, 512 cells per chromosome.
  * stress:  one 250 Mbp chromosome, 1 kb bins, 5 Mbp band, 40 LEFs/Mbp, 12 000 barriers with
             occupancy U(0.5, 0.95), 2048 cells.
A BED record [start, end) places its barrier at (start + end + 1) / 2 (genome.cpp:194-210).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class BedRecord:
    chrom: str
    start: int
    end: int
    score: float
    strand: str


@dataclass
class Workload:
    name: str
    chroms: list  # [(name, size)]
    barriers: list  # [BedRecord]
    options: dict = field(default_factory=dict)  # CLI options given explicitly

    def barriers_of(self, chrom: str):
        return [b for b in self.barriers if b.chrom == chrom]


def _uniform_barriers(rng, chrom, size, n, score_fn):
    starts = np.sort(rng.choice(size - 1, size=n, replace=False))
    strands = np.where(rng.random(n) < 0.5, "+", "-")
    scores = score_fn(n)
    return [BedRecord(chrom, int(s), int(s) + 1, float(sc), str(st))
            for s, st, sc in zip(starts, strands, scores)]


def small(seed: int = 0, ncells: int = 64) -> Workload:
    rng = np.random.default_rng(seed)
    size = 20_000_000
    bars = _uniform_barriers(rng, "chrS", size, 200, lambda n: np.zeros(n))
    return Workload(
        "small-20Mbp-5kb-64cells",
        [("chrS", size)],
        bars,
        dict(bin_size=5_000, diagonal_width=3_000_000, number_of_lefs_per_mbp=20.0,
             avg_lef_processivity=300_000, num_cells=ncells, target_contact_density=1.0,
             seed=seed, extrusion_barrier_occupancy=0.825, barrier_not_occupied_stp=0.7),
    )


def genome(seed: int = 0, ncells: int = 512, nchroms: int = 24) -> Workload:
    rng = np.random.default_rng(seed + 1)
    lengths = np.exp(rng.uniform(np.log(40e6), np.log(250e6), nchroms)).astype(np.int64)
    chroms = [(f"chr{i + 1}", int(L)) for i, L in enumerate(lengths)]
    bars = []
    for name, L in chroms:
        # clustered Poisson process: cluster centres at rate 1/(50 kb * E[size]), E[size] = 2
        ncl = rng.poisson(L / 100_000)
        centres = rng.uniform(0, L - 1, ncl)
        sizes = rng.geometric(0.5, ncl)
        pos = []
        for c, k in zip(centres, sizes):
            off = rng.uniform(-5_000, 5_000, k)
            pos.extend(np.clip(c + off, 0, L - 2).astype(np.int64).tolist())
        pos = np.unique(np.asarray(pos, dtype=np.int64))
        strands = np.where(rng.random(len(pos)) < 0.5, "+", "-")
        occ = rng.beta(8, 2, len(pos))
        occ = np.clip(occ, 1e-6, 1.0)
        bars.extend(BedRecord(name, int(p), int(p) + 1, float(o), str(s))
                    for p, s, o in zip(pos, strands, occ))
    return Workload(
        f"genome-{nchroms}chroms-5kb-{ncells}cells",
        chroms,
        bars,
        dict(bin_size=5_000, diagonal_width=3_000_000, number_of_lefs_per_mbp=20.0,
             num_cells=ncells, target_contact_density=1.0, seed=seed),
    )


def stress(seed: int = 0, ncells: int = 2048) -> Workload:
    rng = np.random.default_rng(seed + 2)
    size = 250_000_000
    bars = _uniform_barriers(rng, "chrL", size, 12_000, lambda n: rng.uniform(0.5, 0.95, n))
    return Workload(
        "stress-250Mbp-1kb-2048cells",
        [("chrL", size)],
        bars,
        dict(bin_size=1_000, diagonal_width=5_000_000, number_of_lefs_per_mbp=40.0,
             avg_lef_processivity=300_000, num_cells=ncells, target_contact_density=1.0,
             seed=seed),
    )
