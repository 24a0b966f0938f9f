"""Seeded synthetic inputs for the simulate path: chrom.sizes entries plus BED6 barrier records.

Each workload is the input a user would hand to `modle simulate` (chromosomes, barrier BED
records and the CLI options that differ from the defaults); nothing here runs the simulation.

  * small:   one 20 Mbp chromosome, 200 uniform barriers, 5 kb bins, 3 Mbp band, 64 cells.
  * genome:  24 chromosomes of log-uniform length in [40, 250] Mbp, clustered barriers with
             Beta(8,2) occupancy, 512 cells per chromosome.
  * stress:  one 250 Mbp chromosome, 1 kb bins, 5 Mbp band, 40 LEFs/Mbp, 12 000 barriers with
             occupancy U(0.5, 0.95), 2048 cells.
  * perturb: `modle perturbate` input: one 30 Mbp chromosome with the genome workload's
             clustered barriers and two feature tracks (1 kb sites, one near barriers, one
             uniform), the reference's defaults otherwise (3 Mbp band, 10 kb deletions, 9 x 9
             blocks).
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
    # feature tracks (--feature-beds) per chromosome: two lists of (start, end, name, strand)
    features: dict = field(default_factory=dict)

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


def _clustered_barriers(rng, name, L):
    """Clustered Poisson process: cluster centres at rate 1/(50 kb * E[size]), E[size] = 2
    barriers within 5 kb of the centre, random strand, occupancy (BED score) Beta(8, 2)."""
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
    return [BedRecord(name, int(p), int(p) + 1, float(o), str(s))
            for p, s, o in zip(pos, strands, occ)]


def genome(seed: int = 0, ncells: int = 512, nchroms: int = 24) -> Workload:
    rng = np.random.default_rng(seed + 1)
    lengths = np.exp(rng.uniform(np.log(40e6), np.log(250e6), nchroms)).astype(np.int64)
    chroms = [(f"chr{i + 1}", int(L)) for i, L in enumerate(lengths)]
    bars = []
    for name, L in chroms:
        bars.extend(_clustered_barriers(rng, name, L))
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


def perturb(seed: int = 0, size: int = 30_000_000, nfeat: int = 120) -> Workload:
    rng = np.random.default_rng(seed + 3)
    name = "chrW"
    bars = _clustered_barriers(rng, name, size)
    # track 1: sites within 20 kb of a random subset of barriers; track 2: uniform sites
    anchor = rng.choice([b.start for b in bars], size=nfeat, replace=False)
    s1 = np.unique(np.clip(anchor + rng.integers(-20_000, 20_000, nfeat), 0, size - 1_000))
    s2 = np.unique(rng.integers(0, size - 1_000, nfeat))
    strand = lambda: "+" if rng.random() < 0.5 else "-"  # noqa: E731
    f1 = [(int(s), int(s) + 1_000, f"a{k}", strand()) for k, s in enumerate(s1.tolist())]
    f2 = [(int(s), int(s) + 1_000, f"b{k}", strand()) for k, s in enumerate(s2.tolist())]
    return Workload(
        f"perturb-{size // 1_000_000}Mbp-5kb",
        [(name, size)],
        bars,
        dict(bin_size=5_000, diagonal_width=3_000_000, number_of_lefs_per_mbp=20.0,
             target_contact_density=1.0, seed=seed),
        {name: [f1, f2]},
    )
