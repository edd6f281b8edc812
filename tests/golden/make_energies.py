"""Writes tests/golden/energies.json: oracle (CPU restatement) Davidson energies
for the synthetic parity cases, converged to residual 1e-8 (energy error
~ residual^2, far below 1e-9 Ha).  The initial vector is the init = 1 vector
(the reference's Randomize with seed 1729, sbd_amd.synth.randomize).  The cases are
densely connected bases: with the reference's diagonal preconditioner,
Davidson stalls on nearly diagonal matrices (for a diagonal H the
preconditioned residual is parallel to the current vector), which is what the
large sparse workloads look like when cut down to test size.  Run from the
repository root:  python tests/golden/make_energies.py
These cases need a few hundred Davidson steps (minutes on 8 cores), so the
energies are committed rather than recomputed inside the GPU tests."""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from sbd_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = {"sci16": ("sci16", 1.0), "sci16open": ("sci16open", 1.0), "fci12": ("fci12", 1.0)}
SEED = 1729

out = {}
for key, (name, scale) in CASES.items():
    w = synth.make_workload(name, scale)
    S = O.System.from_workload(w)
    hii = S.hdiag()
    W0 = synth.randomize(S.N, SEED)
    W, hist = S.davidson(hii, W0, 100, 10, 1e-8)
    E = S.energy(hii, W)
    out[key] = {"workload": name, "scale": scale, "N": int(S.N), "energy": E, "init": 1,
                "seed": SEED, "steps": int(len(hist)), "residual": float(hist[-1][1]),
                "hii_min": float(hii.min())}
    print(key, out[key], flush=True)
    with open(os.path.join(ROOT, "tests", "golden", "energies.json"), "w") as f:
        json.dump(out, f, indent=1)
