"""Serving engine (sbd_server_*): continuous batching of independent solves.

The property the engine is built around: every solve computes the same way
whatever else is in the batch (the same workgroups and the same fixed-order
partial sums as a launch of its own), so its energy is bit-identical

  * run alone on a server,
  * run in a mix with other solves that are admitted and retired around it,
  * and run by the plain solver MultTPB.diag (sbd_diag) with the same options.

Small synthetic requests of uneven norb, electron count and string count.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[serving] {msg}", flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch
    torch.cuda.set_device(0)
    return torch


@pytest.fixture(scope="module")
def requests_():
    from sbd_amd import synth
    return synth.serving_requests(8, 808, norb=(10, 16), nelec=(3, 5), strings=(40, 300))


def opts():
    # up to 480 steps: most of these sparse requests converge to 1e-8, a few do
    # not; agreement is checked either way
    import sbd_amd
    return sbd_amd.SBD(max_it=60, max_nb=8, eps=1e-8, init=1, seed=99)


def test_server_alone_mix_and_plain_agree(torch, requests_):
    import sbd_amd
    say("plain solves")
    plain = []
    for r in requests_:
        m = sbd_amd.MultTPB(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"])
        E, _, rep = m.diag(opts())
        m.close()
        plain.append((E, rep["steps"], rep["converged"]))
    say("each request alone on a server")
    alone = []
    for r in requests_:
        sv = sbd_amd.Server(16 << 30, max_active=4)
        rid = sv.submit(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"], sbd=opts())
        alone.append(sv.wait(rid))
        sv.close()
    say("all requests in one mix, fewer slots than requests")
    sv = sbd_amd.Server(16 << 30, max_active=3)
    ids = [sv.submit(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"], sbd=opts())
           for r in requests_]
    mix = [sv.wait(i) for i in ids]
    st = sv.stats()
    sv.close()
    assert st["peak_active"] >= 2, st
    assert sum(p[2] for p in plain) >= len(plain) // 2
    for q, (p, a, m) in enumerate(zip(plain, alone, mix)):
        assert a["converged"] == m["converged"] == p[2]
        assert a["energy"] == p[0], (q, a["energy"], p[0])
        assert m["energy"] == a["energy"], (q, m["energy"], a["energy"])
        assert a["steps"] == m["steps"] == p[1], (q, a["steps"], m["steps"], p[1])
        assert m["t_submit"] <= m["t_admit"] <= m["t_done"]


def test_server_memory_admission(torch, requests_):
    """With a budget that fits one request at a time the solves still all
    finish (admitted one after another) and give the same energies."""
    import sbd_amd
    rs = requests_[:3]
    sv = sbd_amd.Server(1, max_active=8)  # every request exceeds the budget alone
    ids = [sv.submit(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"], sbd=opts())
           for r in rs]
    res = [sv.wait(i) for i in ids]
    st = sv.stats()
    sv.close()
    assert st["peak_active"] == 1, st
    sv = sbd_amd.Server(16 << 30, max_active=8)
    ids = [sv.submit(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"], sbd=opts())
           for r in rs]
    res2 = [sv.wait(i) for i in ids]
    sv.close()
    for a, b in zip(res, res2):
        assert a["energy"] == b["energy"]
    # solves were serialised: each admitted after the previous finished
    for p, q in zip(res, res[1:]):
        assert q["t_admit"] >= p["t_done"]


def test_server_rejects_bad_requests(torch, requests_):
    import sbd_amd
    r = requests_[0]
    sv = sbd_amd.Server(16 << 30, max_active=2)
    with pytest.raises(sbd_amd.SbdError):
        sv.submit(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"],
                  sbd=sbd_amd.SBD(max_nb=0))
    for bad in (dict(precision=2), dict(method=2), dict(init=2), dict(max_it=0)):
        with pytest.raises(sbd_amd.SbdError):
            sv.submit(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"],
                      sbd=sbd_amd.SBD(**bad))
    with pytest.raises(sbd_amd.SbdError):
        sv.submit(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"],
                  sbd=opts(), nroots=sbd_amd.MAX_ROOTS + 1)
    sv.close()


def test_server_sjf_admission(torch):
    """policy "sjf": with room for one solve at a time, every request admitted
    after the first is the smallest one still queued (same norb and electron
    count, string counts 40..200, so the footprint grows with the count), and
    the energies equal those of arrival-order admission."""
    import sbd_amd
    from sbd_amd import synth
    I0, h, eri = synth.make_integrals(12, 1212, 24)
    counts = [120, 40, 200, 80, 160]  # submission order
    rs = [dict(norb=12, I0=I0, h=h, eri=eri, a=synth.random_strings(12, 4, c, 30 + c),
               b=synth.random_strings(12, 4, c, 60 + c)) for c in counts]
    results = {}
    for policy in ("fifo", "sjf"):
        sv = sbd_amd.Server(1, max_active=8, policy=policy)
        ids = [sv.submit(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"],
                         sbd=opts()) for r in rs]
        results[policy] = [sv.wait(i) for i in ids]
        sv.close()
    for a, b in zip(results["fifo"], results["sjf"]):
        assert a["energy"] == b["energy"] and a["steps"] == b["steps"]
    res = results["sjf"]
    order = sorted(range(len(rs)), key=lambda q: res[q]["t_admit"])
    say(f"sjf admission order by string count: {[counts[q] for q in order]}")
    for i in range(1, len(order)):
        assert counts[order[i]] == min(counts[q] for q in order[i:]), order
    with pytest.raises(sbd_amd.SbdError):
        sbd_amd.Server(1, policy="lifo")


def test_server_config4_shape_sparse_bases(torch):
    """The serving workload's distribution (norb 16-40, n 4-10, string counts
    log-uniform) at string counts 60-1,200: sparse bases, where the solves take
    the per-element sigma kernel.  Alone, in the mix and plain agree bit for bit."""
    import sbd_amd
    from sbd_amd import synth
    rs = synth.serving_requests(10, 4040, norb=(16, 40), nelec=(4, 10), strings=(60, 1200))
    o = sbd_amd.SBD(max_it=20, max_nb=8, eps=1e-7, init=0)
    say("config-4 shape: plain")
    plain = []
    for r in rs:
        m = sbd_amd.MultTPB(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"])
        E, _, rep = m.diag(o)
        m.close()
        plain.append((E, rep["steps"]))
    say("config-4 shape: alone")
    alone = []
    for r in rs:
        sv = sbd_amd.Server(16 << 30, max_active=2)
        alone.append(sv.wait(sv.submit(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"],
                                       eri=r["eri"], sbd=o)))
        sv.close()
    say("config-4 shape: mix")
    sv = sbd_amd.Server(16 << 30, max_active=6, policy="sjf")
    ids = [sv.submit(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"], sbd=o)
           for r in rs]
    mix = [sv.wait(i) for i in ids]
    st = sv.stats()
    sv.close()
    assert st["peak_active"] >= 2, st
    for q, (p, a, m) in enumerate(zip(plain, alone, mix)):
        assert a["energy"] == p[0] and m["energy"] == p[0], (q, p[0], a["energy"], m["energy"])
        assert a["steps"] == m["steps"] == p[1], q


def test_server_fp32_and_multiroot_requests(torch):
    """Requests outside the batched form (FP32 Davidson + FP64 sweep, and the
    block Davidson for several roots, FP64 and FP32 vectors; an open-shell
    basis as in the multi-root serving workload) are served whole beside
    batched single-root solves, and equal MultTPB.diag / diag_roots with the
    same options bit for bit; each stays so when served alone."""
    import sbd_amd
    from sbd_amd import synth
    I0, h, eri = synth.make_integrals(14, 1414, 28)
    base = dict(norb=14, I0=I0, h=h, eri=eri)
    reqs = [
        (dict(base, a=synth.random_strings(14, 5, 600, 11), b=synth.random_strings(14, 4, 500, 12)),
         sbd_amd.SBD(max_it=40, max_nb=8, eps=1e-8, precision=1, fp32_eps=1e-4), 1),
        (dict(base, a=synth.random_strings(14, 5, 700, 13), b=synth.random_strings(14, 4, 400, 14)),
         sbd_amd.SBD(max_it=40, max_nb=24, eps=1e-6), 4),
        (dict(base, a=synth.random_strings(14, 5, 500, 15), b=synth.random_strings(14, 4, 450, 16)),
         sbd_amd.SBD(max_it=40, max_nb=24, eps=1e-6, precision=1), 4),
        (dict(base, a=synth.random_strings(14, 5, 300, 17), b=synth.random_strings(14, 4, 300, 18)),
         sbd_amd.SBD(max_it=40, max_nb=8, eps=1e-8), 1),
    ]
    say("fp32 / multi-root: plain")
    plain = []
    for r, o, nr in reqs:
        m = sbd_amd.MultTPB(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"])
        if nr == 1:
            E, _, rep = m.diag(o)
            plain.append(([E], rep))
        else:
            E, res, rep = m.diag_roots(o, nr)
            plain.append((list(E), rep))
        m.close()
    say("fp32 / multi-root: served together, then alone")
    sv = sbd_amd.Server(16 << 30, max_active=4)
    ids = [sv.submit(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"], sbd=o,
                     nroots=nr) for r, o, nr in reqs]
    mix = [sv.wait(i) for i in ids]
    sv.close()
    alone = []
    for r, o, nr in reqs:
        sv = sbd_amd.Server(16 << 30, max_active=1)
        alone.append(sv.wait(sv.submit(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"],
                                       eri=r["eri"], sbd=o, nroots=nr)))
        sv.close()
    for q, ((Ep, rep), (r, o, nr), m, a) in enumerate(zip(plain, reqs, mix, alone)):
        assert m["nroots"] == nr and len(m["energies"]) == nr
        assert list(m["energies"]) == Ep, (q, m["energies"], Ep)
        assert list(a["energies"]) == Ep, (q, a["energies"], Ep)
        assert m["energy"] == Ep[0] and m["steps"] == rep["steps"], q
        if o.precision == 1 and nr == 1:
            assert m["steps_fp32"] == rep["steps_fp32"] > 0
