"""Wavefunction files (SaveWavefunction / LoadWavefunction, chemistry/tpb/
restart.h:20-65, 414-657) and the warm start of diag (sbdiag.h:339-347).

The file layout is pinned byte for byte against a file built here with struct
from the layout restart.h writes (size_t na, nb, words per string; strings as
bit_length-bit det_vector words; W as doubles; name = prefix + "000000").
Loading into a different basis keeps exactly the coefficients whose alpha and
beta strings are both present, at their new positions, normalised."""
import os
import struct
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


def words(s, L, bl):
    wps = (L + bl - 1) // bl
    return [(int(s) >> (k * bl)) & ((1 << bl) - 1) for k in range(wps)]


def file_bytes(a, b, W, L, bl):
    wps = (L + bl - 1) // bl
    out = struct.pack("<3Q", len(a), len(b), wps)
    for s in list(a) + list(b):
        out += struct.pack(f"<{wps}Q", *words(s, L, bl))
    out += np.ascontiguousarray(W, dtype="<f8").tobytes()
    return out


def test_save_layout_and_reload(torch, tmp_path):
    import sbd_amd
    from sbd_amd import synth
    say("wavefile layout")
    L = 12
    _, h, eri = synth.make_integrals(L, 3, 2 * L)
    a = synth.all_strings(L, 3)[:40]
    b = synth.all_strings(L, 2)[:30]
    m = sbd_amd.MultTPB(L, a, b, I0=0.0, h=h, eri=eri)
    W = synth.random_init_vector(m.N, 8)
    for bl in (20, 5):
        prefix = str(tmp_path / f"wf_bl{bl}_")
        m.save_wavefunction(prefix, W, bl)
        got = open(prefix + "000000", "rb").read()
        assert got == file_bytes(a, b, W, L, bl)
        W2, placed = m.load_wavefunction(prefix, bl)
        assert placed == m.N
        assert np.abs(W2 - W / np.linalg.norm(W)).max() < 1e-15
    m.close()


def test_load_into_other_basis(torch, tmp_path):
    """Saved basis: alpha strings a0 (every other of a list), beta b0; loading
    basis: alpha a1, beta b1 overlapping partly.  Coefficients move to the new
    positions, the others are dropped, the result is normalised."""
    import sbd_amd
    from sbd_amd import synth
    say("wavefile basis change")
    L = 10
    _, h, eri = synth.make_integrals(L, 4, 2 * L)
    sa = synth.all_strings(L, 3)
    sb = synth.all_strings(L, 3)
    a0, b0 = sa[0:60:2], sb[10:50]
    a1, b1 = sa[20:80], sb[0:30]
    rng = np.random.default_rng(2)
    W0 = rng.standard_normal(len(a0) * len(b0))
    path = str(tmp_path / "wf_")
    open(path + "000000", "wb").write(file_bytes(a0, b0, W0, L, 20))
    m = sbd_amd.MultTPB(L, a1, b1, I0=0.0, h=h, eri=eri)
    W, placed = m.load_wavefunction(path, 20)
    exp = np.zeros((len(a1), len(b1)))
    pa = {int(s): i for i, s in enumerate(a1)}
    pb = {int(s): i for i, s in enumerate(b1)}
    n = 0
    for i, s in enumerate(a0):
        for j, t in enumerate(b0):
            if int(s) in pa and int(t) in pb:
                exp[pa[int(s)], pb[int(t)]] = W0[i * len(b0) + j]
                n += 1
    assert placed == n > 0
    exp = exp.reshape(-1) / np.linalg.norm(exp)
    assert np.abs(W - exp).max() < 1e-15
    with pytest.raises(sbd_amd.SbdError):
        m.load_wavefunction(str(tmp_path / "missing_"), 20)
    m.close()


def test_warm_start_h2o(torch, tmp_path):
    """diag saves its vector (savename); a second diag loading it (loadname)
    starts converged: same energy within 1e-9 Ha in far fewer steps."""
    import sbd_amd
    from oracle import oracle as O
    say("warm start h2o")
    hdr, vals, idx = O.load_fcidump(os.path.join(GOLDEN, "h2o", "fcidump.txt.gz"))
    L = int(hdr["NORB"])
    a = O.load_alpha_dets(os.path.join(GOLDEN, "h2o", "h2o-1em3-alpha.txt.gz"), L)
    fd = sbd_amd.FCIDump({"NORB": str(L)}, vals, idx)
    opt = sbd_amd.SBD(max_it=20, max_nb=10, eps=1e-8)
    prefix = str(tmp_path / "h2o_")
    r1 = sbd_amd.diag(opt, fd, a, a, savename=prefix)
    assert os.path.exists(prefix + "000000")
    m = sbd_amd.MultTPB(L, a, a, fcidump=fd)
    W0, placed = m.load_wavefunction(prefix, opt.bit_length)
    E_cold, _, rc = m.diag(opt)
    E_warm, _, rw = m.diag(opt, W_init=W0)
    m.close()
    assert rc["converged"] == 1 and rw["converged"] == 1
    assert abs(E_warm - E_cold) < 1e-9 and abs(r1.energy - E_cold) < 1e-9
    assert rw["steps"] < rc["steps"] // 3, (rw["steps"], rc["steps"])
    r2 = sbd_amd.diag(opt, fd, a, a, loadname=prefix)
    assert abs(r2.energy - E_cold) < 1e-9


def test_initial_determinant_and_matrix_form_h2o(torch, tmp_path):
    """diag from an explicit initial determinant (initial_*determinant_
    bitstring, sbdiag.h:303-335; BasisInitVectorFromDeterminants,
    davidson.h:89-127): from the Hartree-Fock string (index 0) it is init 0,
    bit for bit; from another determinant it equals the solve started from
    that unit vector.  The reference's errors for a missing string, a beta
    string without an alpha one and a string present twice.  SaveMatrixFormWF
    (dump_matrix_form_wf): the .bin file holds the solve's W."""
    import sbd_amd
    from oracle import oracle as O
    say("initial determinant h2o")
    hdr, vals, idx = O.load_fcidump(os.path.join(GOLDEN, "h2o", "fcidump.txt.gz"))
    L = int(hdr["NORB"])
    a = O.load_alpha_dets(os.path.join(GOLDEN, "h2o", "h2o-1em3-alpha.txt.gz"), L)
    fd = sbd_amd.FCIDump({"NORB": str(L)}, vals, idx)
    bits = lambda s: format(int(s), f"0{L}b")
    opt = sbd_amd.SBD(max_it=20, max_nb=10, eps=1e-8)
    r0 = sbd_amd.diag(opt, fd, a, a)
    hf = sbd_amd.SBD(max_it=20, max_nb=10, eps=1e-8, initial_adeterminant_bitstring=bits(a[0]),
                     dump_matrix_form_wf=str(tmp_path / "h2o_wf.bin"))
    r1 = sbd_amd.diag(hf, fd, a, a)
    assert r1.energy == r0.energy
    Wf = np.fromfile(str(tmp_path / "h2o_wf.bin"), dtype="<f8")
    assert Wf.size == len(a) ** 2 and np.array_equal(Wf, r1.W)
    # another determinant: alpha string 3, beta string 5
    m = sbd_amd.MultTPB(L, a, a, fcidump=fd)
    e = np.zeros(m.N)
    e[3 * m.nb + 5] = 1.0
    E_unit, _, _ = m.diag(opt, W_init=e)
    E_det, _, _ = m.diag(opt, start_det=(a[3], a[5]))
    m.close()
    # (this start may lie in another symmetry sector than the ground state,
    # where the solve stays: only the variational bound holds)
    assert E_det == E_unit and E_det >= r0.energy - 1e-9
    od = sbd_amd.SBD(max_it=20, max_nb=10, eps=1e-8, initial_adeterminant_bitstring=bits(a[3]),
                     initial_bdeterminant_bitstring=bits(a[5]))
    assert sbd_amd.diag(od, fd, a, a).energy == E_det
    say("initial determinant errors")
    missing = bits(sum(1 << p for p in range(L - 5, L)))
    for sa, sb, msg in (("", bits(a[0]), "requires an initial alpha"),
                        (missing, "", "initial alpha determinant is not present"),
                        (bits(a[0]), missing, "initial beta determinant is not present"),
                        (bits(a[0])[1:], "", "length does not match")):
        bad = sbd_amd.SBD(max_it=2, initial_adeterminant_bitstring=sa,
                          initial_bdeterminant_bitstring=sb)
        with pytest.raises(ValueError, match=msg):
            sbd_amd.diag(bad, fd, a, a)
    dup = np.concatenate([a, a[:1]])
    with pytest.raises(ValueError, match="occurs more than once"):
        sbd_amd.diag(sbd_amd.SBD(initial_adeterminant_bitstring=bits(a[0])), fd, dup, a)
