"""Writes tests/golden/roots.json: the lowest few eigenvalues of a case's H,
found by scipy's ARPACK Lanczos (scipy.sparse.linalg.eigsh) on a linear
operator that applies the oracle's sigma, with the residual norm of each
returned eigenpair.  These pin the multi-root (block Davidson) solves on bases
too large for a dense eigensolver.  Run from the repository root:
    python tests/golden/make_roots.py [case ...]"""
import json
import os
import sys

import numpy as np
from scipy.sparse.linalg import LinearOperator, eigsh

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from sbd_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = {"sci48openhf": ("sci48openhf", 1.0, 4)}

path = os.path.join(ROOT, "tests", "golden", "roots.json")
data = json.load(open(path)) if os.path.exists(path) else {}
want = sys.argv[1:] or list(CASES)
for key in want:
    name, scale, k = CASES[key]
    w = synth.make_workload(name, scale)
    S = O.System.from_workload(w)
    hii = S.hdiag()
    n_mv = [0]

    def mv(x):
        n_mv[0] += 1
        return S.sigma(hii, np.ascontiguousarray(x.ravel(), dtype=np.float64))

    op = LinearOperator((S.N, S.N), matvec=mv, dtype=np.float64)
    e, V = eigsh(op, k=k, which="SA", v0=synth.randomize(S.N, 7), tol=1e-12, ncv=40,
                 maxiter=5000)
    order = np.argsort(e)
    e, V = e[order], V[:, order]
    res = [float(np.linalg.norm(mv(V[:, j]) - e[j] * V[:, j])) for j in range(k)]
    data[key] = {"workload": name, "scale": scale, "N": int(S.N), "nroots": k,
                 "roots": [float(x) for x in e], "residuals": res, "matvecs": n_mv[0],
                 "hii_min": float(hii.min())}
    print(key, data[key], flush=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
