"""Writes tests/golden/energies.json: oracle (CPU restatement) Davidson energies
for the synthetic parity cases, converged to residual 1e-8 (energy error
~ residual^2, far below 1e-9 Ha).  The initial vector is the init = 1 vector
(the reference's Randomize with seed 1729, sbd_amd.synth.randomize).  The cases are
densely connected bases: with the reference's diagonal preconditioner,
Davidson stalls on nearly diagonal matrices (for a diagonal H the
preconditioned residual is parallel to the current vector), which is what the
large sparse workloads look like when cut down to test size.  Run from the
repository root:  python tests/golden/make_energies.py [case ...]
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

# name: (workload, scale, init); init 1 = the seeded Randomize start, init 0 =
# the Hartree-Fock determinant (the connected norb-58 basis, whose lowest state
# a random start reaches only slowly).  A workload "fcidump:<dir>/<alpha file>"
# is a molecule under tests/golden (the FCIDUMP and alpha strings the reference
# ships; alpha strings used for beta too, as its examples do).
CASES = {"sci16": ("sci16", 1.0, 1), "sci16open": ("sci16open", 1.0, 1), "fci12": ("fci12", 1.0, 1),
         "sci58hf": ("sci58hf", 1.0, 0), "fe4s4": ("fcidump:fe4s4/alpha.txt.gz", 1.0, 0)}
SEED = 1729


def system_of(name, scale):
    """The oracle's System for a case's workload name (both golden scripts)."""
    if name.startswith("fcidump:"):
        mol, alpha = name[len("fcidump:"):].split("/")
        hdr, vals, idx = O.load_fcidump(os.path.join(ROOT, "tests", "golden", mol, "fcidump.txt.gz"))
        L, _, I0, I1, I2, Dm, Xm = O.setup_integrals(hdr, vals, idx)
        a = O.load_alpha_dets(os.path.join(ROOT, "tests", "golden", mol, alpha), L)
        return O.System(L, a, a, I0, I1, I2, Dm, Xm)
    return O.System.from_workload(synth.make_workload(name, scale))


def main(only):
    """(Re)computes the named cases (every case when none is named)."""
    out = {}
    path = os.path.join(ROOT, "tests", "golden", "energies.json")
    if os.path.exists(path):
        out = json.load(open(path))
    for key, (name, scale, init) in CASES.items():
        if only and key not in only:
            continue
        S = system_of(name, scale)
        hii = S.hdiag()
        if init == 1:
            W0 = synth.randomize(S.N, SEED)
        else:
            W0 = np.zeros(S.N)
            W0[0] = 1.0
        W, hist = S.davidson(hii, W0, 100, 10, 1e-8)
        E = S.energy(hii, W)
        prev = out.get(key, {})
        out[key] = {"workload": name, "scale": scale, "N": int(S.N), "energy": E, "init": init,
                    "seed": SEED, "steps": int(len(hist)), "residual": float(hist[-1][1]),
                    "hii_min": float(hii.min())}
        for k in ("energy_eigsh", "eigsh_matvecs"):  # written by make_eigsh.py
            if k in prev:
                out[key][k] = prev[k]
        print(key, out[key], flush=True)
        with open(path, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:])
