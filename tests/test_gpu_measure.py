"""GPU parity of the measurements made after the solve (sbd::tpb::diag,
chemistry/tpb/sbdiag.h:666-777) through the C ABI against the oracle:
string weights (rdmat.h:11-58, 136-182) and occupations (occupation.h:10-95)
within 1e-13 (the sums are ordered differently), carry-over strings (rdmat.h:
60-134) identical; at full size, the weights sum to |W|^2."""
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


@pytest.mark.parametrize("name,scale", [("sci16", 1.0), ("sci16open", 1.0), ("sci36", 0.02)])
def test_measurements_match_oracle(torch, name, scale):
    import sbd_amd
    from oracle import oracle as O
    from sbd_amd import synth
    say(f"measurements {name} x{scale}")
    w = synth.make_workload(name, scale)
    ref = O.System.from_workload(w)
    m = sbd_amd.MultTPB(w["norb"], w["a"], w["b"], I0=w["I0"], h=w["h"], eri=w["eri"])
    # a converged ground state (weights concentrated on few strings) and a
    # random vector (weights spread out)
    E, W1, _ = m.diag(sbd_amd.SBD(max_it=50, max_nb=10, eps=1e-6), want_vector=True)
    W2 = synth.random_init_vector(m.N, 99)
    for W in (W1, W2):
        Wd = torch.tensor(W, device="cuda")
        wa, wb = m.string_weights(Wd)
        ra, rb = ref.weights(W)
        assert np.abs(wa - ra).max() < 1e-13 and np.abs(wb - rb).max() < 1e-13
        occ = m.occupation(Wd)
        rocc = ref.occupation(W)
        assert np.abs(occ - rocc).max() < 1e-13
        for thr, ratio in ((0.01, 0.0), (1e-4, 0.0), (0.0, 0.3)):
            ca, cb, tr = m.carryover(Wd, thr, ratio)
            ia, tra = ref.carryover(W, 0, thr, ratio)
            ib, trb = ref.carryover(W, 1, thr, ratio)
            assert np.array_equal(ca, w["a"][ia]) and np.array_equal(cb, w["b"][ib])
            assert abs(tr[0] - tra) < 1e-13 and abs(tr[1] - trb) < 1e-13
    m.close()


def test_diag_outputs_h2o(torch):
    """The diag mirror on H2O 1e-3: occupations sum to the electron counts (5
    and 5) and lie in [0, 1]; the carry-over strings start with the dominant
    (Hartree-Fock) string and carry at least 1 - threshold of the weight."""
    import sbd_amd
    from oracle import oracle as O
    say("diag outputs h2o")
    hdr, vals, idx = O.load_fcidump(os.path.join(GOLDEN, "h2o", "fcidump.txt.gz"))
    L = int(hdr["NORB"])
    a = O.load_alpha_dets(os.path.join(GOLDEN, "h2o", "h2o-1em3-alpha.txt.gz"), L)
    fd = sbd_amd.FCIDump({"NORB": str(L)}, vals, idx)
    r = sbd_amd.diag(sbd_amd.SBD(max_it=20, max_nb=10, eps=1e-8, carryover_type=1,
                                 threshold=0.01), fd, a, a)
    assert abs(r.energy - (-76.23594663)) < 6e-9
    assert abs(r.density[0::2].sum() - 5.0) < 1e-10 and abs(r.density[1::2].sum() - 5.0) < 1e-10
    assert r.density.min() > -1e-12 and r.density.max() < 1.0 + 1e-12
    assert r.co_adet[0] == a[0] and r.co_bdet[0] == a[0]
    assert 1 <= len(r.co_adet) <= len(a)


def test_weights_full_size(torch):
    import sbd_amd
    from sbd_amd import synth
    say("weights sci36 x0.25")
    w = synth.make_workload("sci36", 0.25)
    m = sbd_amd.MultTPB(w["norb"], w["a"], w["b"], I0=w["I0"], h=w["h"], eri=w["eri"])
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    x = torch.randn(m.N, device="cuda", dtype=torch.float64, generator=g)
    n2 = m.inner(x, x)
    wa, wb = m.string_weights(x)
    assert abs(wa.sum() - n2) < 1e-12 * n2 and abs(wb.sum() - n2) < 1e-12 * n2
    occ = m.occupation(x)
    assert abs(occ[0::2].sum() - 8 * n2) < 1e-11 * n2
    m.close()


def test_singles_extension_amplitude_form(torch):
    """SinglesExtendHalfdets (extend.h:14-185) on a converged sci16 vector:
    the strings of the determinants with W^2 > cutoff, extended by singles,
    against a numpy selection and the string form; the weight above the
    cutoff within 1e-13."""
    import sbd_amd
    from sbd_amd import synth
    say("singles extension sci16")
    w = synth.make_workload("sci16", 1.0)
    m = sbd_amd.MultTPB(w["norb"], w["a"], w["b"], I0=w["I0"], h=w["h"], eri=w["eri"])
    _, W, _ = m.diag(sbd_amd.SBD(max_it=50, max_nb=10, eps=1e-6), want_vector=True)
    Wd = torch.tensor(W, device="cuda")
    for cut in (1e-3, 1e-5):
        ea, eb, tw = m.singles_extend(Wd, cut)
        W2 = (W * W).reshape(m.na, m.nb)
        sel = W2 > cut
        ka = np.asarray(w["a"])[sel.any(axis=1)]
        kb = np.asarray(w["b"])[sel.any(axis=0)]
        assert np.array_equal(ea, sbd_amd.singles_extend_strings(ka, m.norb))
        assert np.array_equal(eb, sbd_amd.singles_extend_strings(kb, m.norb))
        assert abs(tw - W2[sel].sum()) < 1e-13
    m.close()
