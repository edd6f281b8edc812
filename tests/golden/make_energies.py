"""Writes tests/golden/energies.json: oracle (CPU restatement) Davidson energies
for the synthetic parity cases, converged to residual 1e-8 (energy error
~ residual^2, far below 1e-9 Ha).  Run from the repository root:
    python tests/golden/make_energies.py
The FCI case needs several hundred Davidson steps (minutes on 8 cores), so its
energy is committed rather than recomputed inside the GPU tests."""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from sbd_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = {"fci12": ("fci12", 1.0), "sci36_small": ("sci36", 0.04),
         "sci48open_small": ("sci48open", 0.02)}

out = {}
for key, (name, scale) in CASES.items():
    w = synth.make_workload(name, scale)
    S = O.System.from_workload(w)
    hii = S.hdiag()
    W0 = np.zeros(S.N)
    W0[0] = 1.0
    W, hist = S.davidson(hii, W0, 100, 10, 1e-8)
    E = S.energy(hii, W)
    out[key] = {"workload": name, "scale": scale, "N": int(S.N), "energy": E,
                "steps": int(len(hist)), "residual": float(hist[-1][1]),
                "hii_min": float(hii.min())}
    print(key, out[key], flush=True)
with open(os.path.join(ROOT, "tests", "golden", "energies.json"), "w") as f:
    json.dump(out, f, indent=1)
