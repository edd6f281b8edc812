"""GPU parity of the full one- and two-particle RDMs (do_rdm != 0; Correlation,
chemistry/tpb/correlation.h:379-643) through the C ABI against the oracle.

* small bases (10 to 48 orbitals, closed and open shell, dense and sparse
  connections, a single beta string): every entry of both matrices within 1e-12 of the oracle (the
  sums are ordered differently, as the reference's atomics order them
  differently from run to run);
* larger bases, where the oracle would take minutes: size-independent
  identities.  The RDM energy I0 + sum h <a+a> + 1/2 sum (pr|qt) Gamma equals
  <W|H|W> from the sigma (the reference's placement of some alpha x beta terms
  in the other alpha x beta block leaves the sum of those two blocks, and so
  this energy, unchanged); the one-particle traces equal the electron counts;
  two runs give identical bits;
* the diag mirror with do_rdm = 1 reads the density off the one-particle
  diagonal (sbdiag.h:724-727), which equals the do_rdm = 0 occupation density.
"""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, has_gpu

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


def _workload(name):
    from sbd_amd import synth
    if name == "l34":
        # 34 orbitals (the 64-orbital accumulator path), sparse selected strings
        I0, h, eri = synth.make_integrals(34, 4242, 68)
        return dict(norb=34, I0=I0, h=h, eri=eri, a=synth.random_strings(34, 3, 300, 11),
                    b=synth.random_strings(34, 2, 200, 12))
    if name == "l48":
        # 48 orbitals: spectator masks reach past bit 40 (unpacked tables)
        I0, h, eri = synth.make_integrals(48, 4848, 96)
        return dict(norb=48, I0=I0, h=h, eri=eri, a=synth.random_strings(48, 2, 400, 13),
                    b=synth.random_strings(48, 2, 300, 14))
    if name == "one_beta":
        # a single beta string: no beta excitations, no alpha x beta terms
        I0, h, eri = synth.make_integrals(10, 1010, 20)
        return dict(norb=10, I0=I0, h=h, eri=eri, a=synth.all_strings(10, 3),
                    b=np.array([0b0000000111], dtype=np.uint64))
    scale = {"fci12": 0.2, "sci16": 1.0, "sci16open": 1.0}[name]
    return synth.make_workload(name, scale)


def _rdm_energy(w, one, two):
    from sbd_amd import synth
    L = w["norb"]
    eri = synth.unpack_eri(w["eri"], L)
    o = one.reshape(2, L, L)
    t = two.reshape(4, L, L, L, L)
    E = w["I0"] + sum(np.einsum("qp,pq->", o[s], w["h"]) for s in range(2))
    return E + 0.5 * sum(np.einsum("trqp,prqt->", t[k], eri) for k in range(4))


@pytest.mark.parametrize("name", ["fci12", "sci16", "sci16open", "l34", "l48", "one_beta"])
def test_correlation_matches_oracle(torch, name):
    import sbd_amd
    from oracle import oracle as O
    from sbd_amd import synth
    say(f"correlation {name}")
    w = _workload(name)
    ref = O.System.from_workload(w)
    m = sbd_amd.MultTPB(w["norb"], w["a"], w["b"], I0=w["I0"], h=w["h"], eri=w["eri"])
    if name != "one_beta":
        assert ref.counts()[0] > 0 and ref.counts()[2] > 0, "needs singles of both spins"
    W = synth.random_init_vector(m.N, 17)
    W /= np.linalg.norm(W)
    one, two = m.correlation(torch.tensor(W, device="cuda"))
    r1, r2 = ref.correlation(W)
    L = w["norb"]
    assert np.abs(one - r1.reshape(2, L * L)).max() < 1e-12
    assert np.abs(two - r2.reshape(4, L ** 4)).max() < 1e-12
    # the alpha x beta blocks are populated (the oracle pins their layout)
    assert np.abs(r2[1]).max() > 1e-6 and np.abs(r2[2]).max() > 1e-6
    n_alpha = bin(int(w["a"][0])).count("1")
    assert abs(np.trace(one[0].reshape(L, L)) - n_alpha) < 1e-12
    m.close()


def test_correlation_identities_sci36(torch):
    import sbd_amd
    from sbd_amd import synth
    say("correlation identities sci16 x8 strings, sci36 x0.25")
    for name, scale in (("sci16", 4.0), ("sci36", 0.25)):
        w = synth.make_workload(name, scale)
        m = sbd_amd.MultTPB(w["norb"], w["a"], w["b"], I0=w["I0"], h=w["h"], eri=w["eri"])
        E, W, _ = m.diag(sbd_amd.SBD(max_it=4, max_nb=6, eps=1e-10), want_vector=True)
        Wd = torch.tensor(W, device="cuda")
        n2 = m.inner(Wd, Wd)
        hii = m.makeQChamDiagTerms()
        HW = torch.zeros_like(Wd)
        m.run(hii, Wd, HW)
        E_sigma = m.inner(Wd, HW) / n2
        one, two = m.correlation(Wd)
        one2, two2 = m.correlation(Wd)
        assert np.array_equal(one, one2) and np.array_equal(two, two2)
        L = w["norb"]
        na_e = bin(int(w["a"][0])).count("1")
        nb_e = bin(int(w["b"][0])).count("1")
        assert abs(np.trace(one[0].reshape(L, L)) - na_e * n2) < 1e-11 * n2
        assert abs(np.trace(one[1].reshape(L, L)) - nb_e * n2) < 1e-11 * n2
        E_rdm = (_rdm_energy(w, one, two) - w["I0"]) / n2 + w["I0"]
        say(f"  {name} x{scale}: E_sigma {E_sigma:.12f} E_rdm {E_rdm:.12f}")
        assert abs(E_rdm - E_sigma) < 1e-10 * abs(E_sigma)
        m.close()


def test_diag_do_rdm_h2o(torch):
    import sbd_amd
    from oracle import oracle as O
    say("diag do_rdm=1 h2o")
    hdr, vals, idx = O.load_fcidump(os.path.join(GOLDEN, "h2o", "fcidump.txt.gz"))
    L = int(hdr["NORB"])
    a = O.load_alpha_dets(os.path.join(GOLDEN, "h2o", "h2o-1em3-alpha.txt.gz"), L)
    fd = sbd_amd.FCIDump({"NORB": str(L)}, vals, idx)
    r0 = sbd_amd.diag(sbd_amd.SBD(max_it=20, max_nb=10, eps=1e-8), fd, a, a)
    r1 = sbd_amd.diag(sbd_amd.SBD(max_it=20, max_nb=10, eps=1e-8, do_rdm=1), fd, a, a)
    assert r0.energy == r1.energy
    assert np.abs(r1.density - r0.density).max() < 1e-13
    assert r1.one_p_rdm.shape == (2, L * L) and r1.two_p_rdm.shape == (4, L ** 4)
    # symmetric one-particle RDM of a real state
    for s in range(2):
        g = r1.one_p_rdm[s].reshape(L, L)
        assert np.abs(g - g.T).max() < 1e-13
