"""Adds to tests/golden/energies.json the lowest eigenvalue of each case's H
found by scipy's ARPACK Lanczos (scipy.sparse.linalg.eigsh) on a linear
operator that applies the oracle's sigma: an energy independent of the
Davidson restatement (the sigma itself is pinned element by element against
second quantisation).  Run from the repository root after make_energies.py:
    python tests/golden/make_eigsh.py [case ...]
(no case names: every case that has no eigsh energy yet)."""
import json
import os
import sys

import numpy as np
from scipy.sparse.linalg import LinearOperator, eigsh

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from sbd_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402,F401
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from make_energies import system_of  # noqa: E402

path = os.path.join(ROOT, "tests", "golden", "energies.json")
data = json.load(open(path))
want = sys.argv[1:] or [k for k, d in data.items() if "energy_eigsh" not in d]
for key, d in data.items():
    if key not in want:
        continue
    S = system_of(d["workload"], d["scale"])
    hii = S.hdiag()
    n_mv = [0]

    def mv(x):
        n_mv[0] += 1
        return S.sigma(hii, np.ascontiguousarray(x.ravel(), dtype=np.float64))

    op = LinearOperator((S.N, S.N), matvec=mv, dtype=np.float64)
    v0 = synth.randomize(S.N, 7)
    e = eigsh(op, k=1, which="SA", v0=v0, tol=1e-13, ncv=40, maxiter=5000,
              return_eigenvectors=False)
    d["energy_eigsh"] = float(e[0])
    d["eigsh_matvecs"] = n_mv[0]
    print(key, d["energy"], d["energy_eigsh"], d["energy"] - d["energy_eigsh"], n_mv[0],
          flush=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
