"""Fingerprints of fast-mode bands: SHA-256 over every chromosome's band and missed count, with
the run's epoch, burn-in and contact totals, for three small seeded workloads. Fast mode's
results depend only on the seeds and the parameters, so these change only when its random
streams or its physics change on purpose; then this script (run on a GPU) rewrites
fast_fingerprints.json and the change is stated in the commit that carries it."""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from modlefork_amd import _lib as L  # noqa: E402
from modlefork_amd import simulation as S  # noqa: E402
from modlefork_amd import workloads as W  # noqa: E402

PATH = os.path.join(HERE, "fast_fingerprints.json")


def cases():
    bypass = W.small(ncells=8)
    bypass.options["probability_of_extrusion_unit_bypass"] = 0.5
    return {"small-16cells": W.small(ncells=16),
            "genome-6chroms-2cells": W.genome(ncells=2, nchroms=6),
            "small-bypass-0.5-8cells": bypass}


def fingerprint(ctx, wl):
    out, st = S.Simulation(wl.options, ctx).run_simulate(wl, L.MODE_FAST)
    h = hashlib.sha256()
    for name in sorted(out):
        band, missed = out[name][0], out[name][1]
        h.update(band.tobytes())
        h.update(str(missed).encode())
    return {"sha256": h.hexdigest(), "epochs": int(st.epochs),
            "burnin_epochs": int(st.burnin_epochs), "contacts": int(st.contacts)}


def main():
    path = sys.argv[1] if len(sys.argv) > 1 else PATH
    with S.Context(0) as ctx:
        res = {name: fingerprint(ctx, wl) for name, wl in cases().items()}
    with open(path, "w") as fh:
        json.dump(res, fh, indent=1)
        fh.write("\n")
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
