"""A small seeded genome for the perturbate / replay tests: one 6 Mbp chromosome with barriers and
two feature tracks, and one 3 Mbp chromosome with barriers but no features (skipped by the plan,
written empty by replay). 5 kb bins, a 1 Mbp band, so windows are 4 Mbp wide (80 LEFs)."""
import numpy as np

from modlefork_amd import perturbate as P
from modlefork_amd import simulation as S
from modlefork_amd import workloads as W

OPTIONS = dict(bin_size=5_000, diagonal_width=1_000_000, number_of_lefs_per_mbp=20.0,
               target_contact_density=0.05, seed=7)
GENOME = [("chrP", 6_000_000), ("chrQ", 3_000_000)]


def barrier_records(name, size, n, rng):
    pos = np.sort(rng.choice(size - 10, n, replace=False) + 5)
    recs = []
    for p in pos.tolist():
        score = 0.0 if rng.random() < 0.2 else float(rng.uniform(0.5, 0.95))
        recs.append(W.BedRecord(name, p - 1, p, score, "+" if rng.random() < 0.5 else "-"))
    # a few records sharing a start (and so a midpoint) with another, in file order
    for p in pos[::9].tolist():
        recs.append(W.BedRecord(name, p - 1, p + 1, 0.7, "-"))
    order = rng.permutation(len(recs))
    return [recs[i] for i in order]


def features(name, size, n, tag, rng):
    starts = np.sort(rng.choice(size // 1000 - 4, n, replace=False) * 1000 + 1000)
    return [P.Feature(name, int(s), int(s) + 2000, f"{tag}{k}", 0.0, "+" if rng.random() < 0.5
                      else "-") for k, s in enumerate(starts.tolist())]


def build(seed=3):
    """(params, ChromInput list, BED records by chromosome)."""
    rng = np.random.default_rng(seed)
    params = S.make_params(OPTIONS)
    chroms, records = [], {}
    for name, size in GENOME:
        recs = barrier_records(name, size, 40 if name == "chrP" else 20, rng)
        records[name] = recs
        imp = S.import_barriers(params, recs, 0, size)
        f1 = features(name, size, 15, "a", rng) if name == "chrP" else []
        f2 = features(name, size, 12, "b", rng) if name == "chrP" else []
        chroms.append(P.ChromInput(name, size, 0, size, imp, f1, f2))
    return params, chroms, records
