"""Writes tests/golden/synth_digests.json: SHA-256 of the seeded synthetic
integrals and string lists for each named workload, so the generator that
defines the benchmark inputs cannot drift.  Run from the repository root."""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from sbd_amd import synth  # noqa: E402

out = {}
for name, scale in [("fci12", 1.0), ("sci36", 1.0), ("sci48open", 1.0), ("sci58", 1.0)]:
    w = synth.make_workload(name, scale)
    out[name] = {
        "scale": scale, "na": int(len(w["a"])), "nb": int(len(w["b"])),
        "h": hashlib.sha256(np.ascontiguousarray(w["h"]).tobytes()).hexdigest(),
        "eri": hashlib.sha256(np.ascontiguousarray(w["eri"]).tobytes()).hexdigest(),
        "a": hashlib.sha256(w["a"].tobytes()).hexdigest(),
        "b": hashlib.sha256(w["b"].tobytes()).hexdigest(),
    }
    print(name, out[name]["na"], out[name]["nb"], flush=True)
with open(os.path.join(ROOT, "tests", "golden", "synth_digests.json"), "w") as f:
    json.dump(out, f, indent=1)
