"""GPU parity of the restarted Lanczos (method 2; chemistry/tpb/lanczos.h:162-327)
through the C ABI against the oracle's restatement (oracle/sbd_oracle.c,
orc_lanczos, itself pinned by tests/test_oracle.py against a dense Krylov
projection and the dense ground state).

Bars: per-step E0 within 1e-10 relative and beta within 1e-8 relative of the
oracle over three passes (the two sum their dot products in different orders,
and Lanczos without reorthogonalisation amplifies those last-bit differences
from step to step); the final vectors within 1e-7; diag with method 3 is
method 2 bit for bit; the H2O energy within 1e-5 Ha of the reference-held one.
"""
import sys

import numpy as np
import pytest

from conftest import has_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as t
    t.cuda.set_device(0)
    return t


@pytest.mark.parametrize("name,scale", [("sci16", 1.0), ("sci16open", 1.0), ("sci36", 0.02)])
def test_lanczos_steps_match_oracle(torch, name, scale):
    import sbd_amd
    from oracle import oracle as O
    from sbd_amd import synth
    say(f"lanczos parity {name} x{scale}")
    w = synth.make_workload(name, scale)
    ref = O.System.from_workload(w)
    m = sbd_amd.MultTPB(w["norb"], w["a"], w["b"], I0=w["I0"], h=w["h"], eri=w["eri"])
    hii = m.hdiag()
    hii_ref = ref.hdiag()
    W0 = synth.random_init_vector(m.N, 1729)
    Wd = torch.tensor(W0, device="cuda")
    E, rep, hist = m.lanczos(hii, Wd, sbd_amd.SBD(max_it=3, max_nb=8, eps=1e-300),
                             history=True)
    Wr, hr = ref.lanczos(hii_ref, W0, 3, 8, 1e-300)
    assert rep["steps"] == len(hr) == 24
    scale_e = np.abs(hr[:, 0]).max()
    assert np.abs(hist[:, 0] - hr[:, 0]).max() < 1e-10 * scale_e
    bmask = hr[:, 2] != 0.0
    assert np.array_equal(hist[:, 1] != 0.0, bmask)
    assert np.abs(hist[bmask, 1] - hr[bmask, 2]).max() < 1e-8 * np.abs(hr[:, 2]).max()
    assert abs(E - hr[-1, 0]) < 1e-10 * scale_e
    assert np.abs(Wd.cpu().numpy() - Wr).max() < 1e-7
    m.close()


def test_diag_lanczos_h2o(torch):
    """diag with method 2 on H2O 1e-3 from the HF determinant reaches the
    reference-held energy (data/h2o/README.md:6, -76.23594663) within 1e-5 Ha
    at eps 1e-7 (the oracle's Lanczos gives -76.2359464697 here); method 3 is
    method 2 bit for bit; Lanczos with precision 1, or an unknown method, is an
    argument error.  (The reference's restarted Lanczos keeps no orthogonality:
    on the synthetic cases it loses it once E0 is converged to ~1e-6 Ha, so
    energies are checked on this case, where it stops cleanly.)"""
    import os
    import sbd_amd
    from conftest import GOLDEN
    from oracle import oracle as O
    say("lanczos h2o 1e-3")
    hdr, vals, idx = O.load_fcidump(os.path.join(GOLDEN, "h2o", "fcidump.txt.gz"))
    L = int(hdr["NORB"])
    a = O.load_alpha_dets(os.path.join(GOLDEN, "h2o", "h2o-1em3-alpha.txt.gz"), L)
    fd = sbd_amd.FCIDump({"NORB": str(L)}, vals, idx)
    m = sbd_amd.MultTPB(L, a, a, fcidump=fd)
    E2, W2, r2 = m.diag(sbd_amd.SBD(method=2, max_it=100, max_nb=10, eps=1e-7),
                        want_vector=True)
    E3, W3, r3 = m.diag(sbd_amd.SBD(method=3, max_it=100, max_nb=10, eps=1e-7),
                        want_vector=True)
    assert r2["converged"] == 1
    assert E2 == E3 and np.array_equal(W2, W3) and r2["steps"] == r3["steps"]
    assert abs(E2 - (-76.23594663)) < 1e-5, E2
    with pytest.raises(sbd_amd.SbdError):
        m.diag(sbd_amd.SBD(method=2, precision=1))
    with pytest.raises(sbd_amd.SbdError):
        m.diag(sbd_amd.SBD(method=4))
    m.close()
