#!/usr/bin/env python3
"""Benchmark of the TPB sigma build (H * v) on one MI355X.

A step is one FP64 sigma build Wb = H Wk over the whole basis of the
workload: norb=36, n_alpha=n_beta=8, 20,000 alpha x 20,000 beta strings
(4.0e8 determinants), seeded synthetic integrals (ERI rank 72).  Inputs are
resident in HBM when the timed region starts.

Printed JSON line:
  value        sigma GDet/s = N / (seconds per sigma)
  roofline     algorithmic bytes of one launch / its average duration, vs 8 TB/s:
               B = 8 (N + nb * P_alpha) + 8 N   (read T row + one T row per alpha
               pair; write W); hii, tables, strings and integrals are not counted
  davidson     time-to-energy @1e-6 Ha: plain FP64 Davidson vs FP32 Davidson +
               FP64 correction, same workload
  cpu_baseline the CPU oracle (restatement of the reference's OpenMP sigma) on
               a bounded sample of bra rows of the same workload
  rdm          wall time of the full one- and two-particle RDMs (do_rdm != 0)
               of a normalised vector on the same workload
  dense        FCI norb 12 (245,025 dets, 1,424 terms/det): sigma GDet/s and
               terms/s, Davidson to residual 1e-8 vs the oracle's energy, and
               the oracle's own sigma and Davidson on the host; the
               reference's C2H2 (100k TrimCI core, 6.19e7 dets) and H2O
               (threshold 1e-6, 1.91e8 dets) inputs solved to residual 1e-6,
               energies against the values the reference's READMEs hold
  precision    norb 58, 10 + 10 electrons, the 60,000 x 60,000 sampled strings
               plus the HF singles (3.66e9 dets): plain FP64 Davidson vs FP32
               Davidson + FP64 sweep, energy error and time to 1e-6 Ha
  serving      a Poisson stream of 256 independent solves (norb 16-40, n 4-10,
               200-8000 strings per spin) through the continuous-batching
               server vs the same stream solved one at a time: p50/p99
               latency and solves/s
  multiroot    32 concurrent open-shell block-Davidson requests (4 roots,
               norb 48, 12 + 10 electrons, the sci48open string counts x 0.1
               with +-30 % jitter, lowest-diagonal start): FP32 vectors with FP64
               Ritz / Gram-Schmidt vs the same requests in FP64, per-root
               energy error, p50/p99 latency, solves/s
"""
import argparse
import gzip
import json
import math
import os
import shutil
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD_DESC = {
    "sci36": "norb=36 na=nb=8, 20000x20000 strings (4.0e8 dets), ERI rank 72",
    "fci12": "norb=12 na=nb=4, FCI 495x495 strings (245,025 dets), ERI rank 24",
    "sci58": "norb=58 na=nb=10, 60000x60000 strings (3.6e9 dets), ERI rank 116",
    "sci48open": "norb=48 na=12 nb=10, 30000x25000 strings (7.5e8 dets), ERI rank 96",
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(w, m, T_dev, Wb_dev, hii_dev, target_s):
    """Time the CPU oracle's sigma on a bounded sample of bra rows; check the
    sampled rows bit for bit against the GPU result."""
    import numpy as np
    from oracle import oracle as O
    t0 = time.time()
    S = O.System.from_workload(w)
    log(f"cpu: oracle tables built in {time.time() - t0:.1f}s")
    nb = S.nb
    T = T_dev.cpu().numpy()
    hii = np.zeros(S.N)
    Wb = np.zeros(S.N)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rows = max(threads, 8)
    S.hdiag_rows(hii, 0, rows)
    t1 = time.time()
    S.sigma_rows(hii, T, Wb, 0, rows)
    dt = time.time() - t1
    # scale the sample to ~target_s of CPU work
    rows2 = int(min(S.na, max(rows, rows * target_s / max(dt, 1e-3))))
    if rows2 > rows:
        S.hdiag_rows(hii, rows, rows2)
        Wb[:] = 0.0
        t1 = time.time()
        S.sigma_rows(hii, T, Wb, 0, rows2)
        dt = time.time() - t1
        rows = rows2
    gpu_rows = Wb_dev[: rows * nb].cpu().numpy()
    gpu_hii = hii_dev[: rows * nb].cpu().numpy()
    exact = bool(np.array_equal(gpu_rows, Wb[: rows * nb]) and
                 np.array_equal(gpu_hii, hii[: rows * nb]))
    dets = rows * nb
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                 if l.startswith("model name")][0]
    except Exception:
        model = "unknown"
    return {"value": dets / dt / 1e9, "unit": "GDet/s", "cores": threads, "kind": "port",
            "sample": f"bra rows 0..{rows - 1} x {nb} beta strings ({dets:.3e} dets) of the "
                      f"same workload, one FP64 sigma, {dt:.1f}s, OpenMP, {model}",
            "sample_bitexact_vs_gpu": exact}


GOLDEN = os.path.join(ROOT, "tests", "golden")


def terms_per_sigma(m, counts):
    """Off-diagonal matrix elements one sigma evaluates: each alpha pair with
    every beta string, each beta pair with every alpha string, and each
    (alpha single, beta single) pair."""
    ca_s, ca_d, cb_s, cb_d = (int(x) for x in counts)
    return m.nb * (ca_s + ca_d) + m.na * (cb_s + cb_d) + ca_s * cb_s


def time_sigma(m, hii, T, steps, warmup=2):
    """Average milliseconds of one sigma launch in T's precision (HIP events on
    the stream the launches go to)."""
    import torch
    Wb = torch.empty_like(T)
    for _ in range(warmup):
        m.run(hii, T, Wb, accumulate=False)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        m.run(hii, T, Wb, accumulate=False)
    e1.record(stream)
    torch.cuda.synchronize()
    del Wb
    return e0.elapsed_time(e1) / steps


def time_to_energy(hist, Eref, tol=1e-6):
    """Seconds until an FP64 Ritz value lies within tol of Eref (history rows
    (E0, |R|, seconds, precision))."""
    for E0, _, t, p in hist:
        if p == 64 and abs(E0 - Eref) <= tol:
            return round(float(t), 4)
    return None


def _plain_copy(tmp, rel):
    out = os.path.join(tmp, os.path.basename(rel)[:-3])
    with gzip.open(os.path.join(GOLDEN, rel), "rb") as f, open(out, "wb") as g:
        shutil.copyfileobj(f, g)
    return out


# the reference's inputs and the energies its READMEs hold (8 decimals)
MOLECULES = [
    ("c2h2_100k", "c2h2/fcidump.txt.gz", "c2h2/alpha-100k.txt.gz", -77.11337996,
     "data/c2h2/README.md:7"),
    ("h2o_1em6", "h2o/fcidump.txt.gz", "h2o/h2o-1em6-alpha.txt.gz", -76.24377593,
     "data/h2o/README.md:9"),
    ("n2_1em7", "n2/fcidump.txt.gz", "n2/1em7-alpha.txt.gz", -109.04874198,
     "data/n2/README.md:14"),
]
# larger inputs, run with --large-molecules (minutes each)
MOLECULES_LARGE = [
    ("h2o_1em7", "h2o/fcidump.txt.gz", "h2o/h2o-1em7-alpha.txt.gz", -76.24377678,
     "data/h2o/README.md:10"),
    ("n2_fci", "n2/fcidump.txt.gz", "n2/full-alpha.txt.gz", -109.04874199,
     "data/n2/README.md:15"),
    ("h2o_1em8", "h2o/fcidump.txt.gz", "h2o/h2o-1em8-alpha.txt.gz", -76.24377680,
     "data/h2o/README.md:11"),
]


def dense_section(args):
    """The dense workload (FCI norb 12) on the GPU and on the host, and the
    reference's larger molecule inputs against their held energies."""
    import numpy as np
    import torch
    import sbd_amd
    from sbd_amd import synth
    out = {}
    w = synth.make_workload("fci12", 1.0)
    m = sbd_amd.MultTPB(w["norb"], w["a"], w["b"], I0=w["I0"], h=w["h"], eri=w["eri"])
    counts = m.table_counts()
    terms = terms_per_sigma(m, counts)
    hii = m.hdiag()
    g = torch.Generator(device="cuda")
    g.manual_seed(12)
    T = torch.randn(m.N, dtype=torch.float64, device="cuda", generator=g)
    ms = time_sigma(m, hii, T, 20)
    gold = json.load(open(os.path.join(GOLDEN, "energies.json")))["fci12"]
    opt = sbd_amd.SBD(max_it=100, max_nb=10, eps=1e-8, init=gold["init"], seed=gold["seed"])
    E, _, rep, hist = m.diag(opt, history=True)
    fci = {"determinants": m.N, "terms_per_det": round(terms / m.N, 1),
           "sigma_ms": round(ms, 4), "sigma_gdet_s": round(m.N / ms * 1e-6, 3),
           "sigma_gterms_s": round(terms / ms * 1e-6, 2),
           "energy": E, "abs_err_vs_oracle": abs(E - gold["energy"]),
           "steps": rep["steps"], "converged": bool(rep["converged"]),
           "t_davidson_s": round(rep["t_davidson"], 4),
           "time_to_energy_s": time_to_energy(hist, gold["energy"]), "eps": 1e-8,
           "bound": "L2/VALU (the alpha x beta pass gathers integrals per term), not HBM"}
    log(f"dense fci12: sigma {ms:.3f} ms ({fci['sigma_gterms_s']} Gterms/s), "
        f"E={E:.12f} |dE|={fci['abs_err_vs_oracle']:.1e} steps={rep['steps']} "
        f"t={rep['t_davidson']:.3f}s")
    if not args.no_cpu:
        # the oracle's sigma and Davidson on the host, same inputs and options
        from oracle import oracle as O
        S = O.System.from_workload(w)
        h = S.hdiag()
        X = T.cpu().numpy()
        t0 = time.time()
        Y = S.sigma(h, X)
        t_sig = time.time() - t0
        exact = bool(np.array_equal(Y, m.sigma(hii, T).cpu().numpy()))
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        fci["cpu_oracle"] = {"sigma_s": round(t_sig, 4), "sigma_gdet_s": round(S.N / t_sig / 1e9, 5),
                             "cores": threads, "kind": "port", "sigma_bitexact_vs_gpu": exact}
        # the oracle's Davidson makes one sigma per step, as the GPU solve did:
        # run it whole when that fits the time allowed, else report the estimate
        est = t_sig * (rep["steps"] + 1)
        if est <= args.cpu_davidson_s:
            W0 = synth.randomize(S.N, gold["seed"])
            t0 = time.time()
            Wc, hc = S.davidson(h, W0, 100, 10, 1e-8)
            t_dav = time.time() - t0
            fci["cpu_oracle"].update({"davidson_s": round(t_dav, 3), "davidson_steps": len(hc),
                                      "davidson_energy": S.energy(h, Wc)})
            log(f"dense fci12 cpu oracle: sigma {t_sig:.3f}s, Davidson {t_dav:.2f}s "
                f"({len(hc)} steps), bit-exact sigma {exact}")
        else:
            fci["cpu_oracle"]["davidson_s_estimate"] = round(est, 1)
            log(f"dense fci12 cpu oracle: sigma {t_sig:.3f}s; Davidson skipped (~{est:.0f}s)")
    out["fci12"] = fci
    m.close()
    del T, hii
    large_names = [x for x in args.large_molecules.split(",") if x]
    known = [x[0] for x in MOLECULES_LARGE]
    for x in large_names:
        if x not in known:
            raise SystemExit(f"--large-molecules: unknown input {x} (one of {', '.join(known)})")
    mols = ([] if args.no_molecules else MOLECULES) + [x for x in MOLECULES_LARGE
                                                       if x[0] in large_names]
    if not mols:
        return out
    with tempfile.TemporaryDirectory() as tmp:
        for name, fd_rel, a_rel, E_held, cite in mols:
            t0 = time.time()
            fd = sbd_amd.LoadFCIDump(_plain_copy(tmp, fd_rel))
            a = sbd_amd.LoadAlphaDets(_plain_copy(tmp, a_rel), fd.norb)
            m = sbd_amd.MultTPB(fd.norb, a, a, fcidump=fd)
            t_setup = time.time() - t0
            counts = m.table_counts()
            terms = terms_per_sigma(m, counts)
            hii = m.hdiag()
            T = torch.randn(m.N, dtype=torch.float64, device="cuda", generator=g)
            ms = time_sigma(m, hii, T, 2, warmup=1)
            del hii
            T32 = T.float()
            del T
            torch.cuda.empty_cache()
            h32 = m.hdiag(torch.float32)
            ms32 = time_sigma(m, h32, T32, 2, warmup=1)
            del T32, h32
            torch.cuda.empty_cache()  # the solve's vectors get the memory torch held
            log(f"dense {name}: N={m.N:.3e} sigma {ms:.2f} ms (FP32 {ms32:.2f} ms)")
            large = (name, fd_rel, a_rel, E_held, cite) in MOLECULES_LARGE
            # the large solves print each Davidson step (the solver's verbose
            # lines, sent to stderr: stdout carries the result line only)
            saved = None
            if large:
                sys.stdout.flush()
                saved = os.dup(1)
                os.dup2(2, 1)
            try:
                E, _, rep = m.diag(sbd_amd.SBD(max_it=60, max_nb=10, eps=1e-6,
                                               verbose=1 if large else 0))
            finally:
                if saved is not None:
                    os.dup2(saved, 1)
                    os.close(saved)
            out[name] = {"determinants": m.N, "norb": fd.norb, "strings": len(a),
                         "terms_per_det": round(terms / m.N, 1), "sigma_ms": round(ms, 3),
                         "sigma_fp32_ms": round(ms32, 3),
                         "sigma_gdet_s": round(m.N / ms * 1e-6, 3),
                         "sigma_gterms_s": round(terms / ms * 1e-6, 2),
                         "energy": E, "energy_held": E_held, "held_at": cite,
                         "abs_err_vs_held": abs(E - E_held), "steps": rep["steps"],
                         "converged": bool(rep["converged"]),
                         "t_davidson_s": round(rep["t_davidson"], 3),
                         "setup_s": round(t_setup, 2), "eps": 1e-6}
            log(f"dense {name}: N={m.N:.3e} sigma {ms:.2f} ms, E={E:.10f} "
                f"(held {E_held}) steps={rep['steps']} t={rep['t_davidson']:.2f}s")
            m.close()
    return out


def precision_section(args):
    """The precision workload at its size: plain FP64 Davidson against FP32
    Davidson + FP64 sweep on 3.66e9 determinants, both from the HF determinant
    (init 0)."""
    import sbd_amd
    from sbd_amd import synth
    t0 = time.time()
    w = synth.make_workload("sci58hf", 100.0)
    m = sbd_amd.MultTPB(w["norb"], w["a"], w["b"], I0=w["I0"], h=w["h"], eri=w["eri"])
    log(f"precision: sci58 strings + HF singles, {len(w['a'])} x {len(w['b'])} = {m.N:.3e} dets, "
        f"set up in {time.time() - t0:.1f}s")
    del w
    res = {}
    for label, prec in (("fp64", 0), ("mixed", 1)):
        opt = sbd_amd.SBD(max_it=args.prec_max_it, max_nb=10, eps=args.prec_eps, init=0,
                          precision=prec)
        E, _, rep, hist = m.diag(opt, history=True)
        res[label] = (E, rep, hist)
        log(f"precision {label}: E={E:.12f} steps={rep['steps']} (fp32 {rep['steps_fp32']}) "
            f"nb {rep['nb_used']}/{rep['nb_used_fp32']} conv={rep['converged']} "
            f"t={rep['t_davidson']:.2f}s")
    Eref = res["fp64"][0]
    out = {"workload": "norb=58 na=nb=10, sci58's 60000 x 60000 sampled strings plus the 480 "
                       "HF singles per spin, ERI rank 116, init 0 (HF start)",
           "determinants": m.N, "eps": args.prec_eps}
    for label, (E, rep, hist) in res.items():
        out[label] = {"energy": E, "abs_err_vs_fp64": abs(E - Eref),
                      "time_to_energy_s": time_to_energy(hist, Eref),
                      "t_davidson_s": round(rep["t_davidson"], 3), "steps": rep["steps"],
                      "steps_fp32": rep["steps_fp32"], "converged": bool(rep["converged"]),
                      "residual": rep["residual"], "subspace_fp64": rep["nb_used"],
                      "subspace_fp32": rep["nb_used_fp32"]}
    t64, tmx = out["fp64"]["time_to_energy_s"], out["mixed"]["time_to_energy_s"]
    out["speedup_mixed_vs_fp64"] = round(t64 / tmx, 3) if t64 and tmx else None
    m.close()
    return out


def multiroot_section(args):
    """The multi-root workload, strings scaled down: concurrent block-Davidson
    requests for the 4 lowest roots (sampled strings as configured, started on the lowest diagonal
    elements), served in
    FP32 (FP64 Ritz and Gram-Schmidt, FP64 finish) and again in FP64; per-root
    energy error of the FP32 results against the FP64 ones."""
    import numpy as np
    import sbd_amd
    from sbd_amd import synth
    reqs = synth.multiroot_requests(args.mr_count, 5151, args.mr_scale, hf=False)
    dets = [len(r["a"]) * len(r["b"]) for r in reqs]
    log(f"multiroot: {len(reqs)} requests, {min(dets):.2e}..{max(dets):.2e} dets")
    out = {"requests": len(reqs), "nroots": 4, "string_scale": args.mr_scale, "jitter": 0.3,
           "determinants_min": min(dets), "determinants_max": max(dets)}
    got = {}
    for label, prec in (("fp32", 1), ("fp64", 0)):
        # init 0: unit vectors on the lowest diagonal elements (from random
        # vectors the diagonal preconditioner returns corrections nearly
        # parallel to the iterate on these nearly diagonal bases)
        opt = sbd_amd.SBD(max_it=args.mr_max_it, max_nb=24, eps=1e-6, init=0, precision=prec,
                          fp32_eps=1e-4)
        sv = sbd_amd.Server(args.serve_mem_gb << 30, max_active=args.serve_slots)
        ids = [sv.submit(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"], sbd=opt,
                         nroots=4) for r in reqs]
        res = []
        for i in ids:
            res.append(sv.wait(i))
            if len(res) % 8 == 0:
                log(f"multiroot {label}: {len(res)} solves done")
        sv.close()
        lat = [r["t_done"] - r["t_submit"] for r in res]
        span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
        got[label] = res
        out[label] = {"solves_per_s": round(len(res) / span, 3), "p50_s": round(pct(lat, 50), 4),
                      "p99_s": round(pct(lat, 99), 4), "makespan_s": round(span, 3),
                      "converged": sum(r["converged"] for r in res),
                      "steps_median": pct([r["steps"] for r in res], 50),
                      "steps_fp32_median": pct([r["steps_fp32"] for r in res], 50)}
        log(f"multiroot {label}: p50 {pct(lat, 50):.2f}s p99 {pct(lat, 99):.2f}s "
            f"{len(res) / span:.2f} solves/s, converged {out[label]['converged']}/{len(res)}")
    err = np.array([[abs(a - b) for a, b in zip(x["energies"], y["energies"])]
                    for x, y in zip(got["fp32"], got["fp64"])])
    out["abs_err_fp32_vs_fp64_per_root_max"] = [float(v) for v in err.max(axis=0)]
    out["abs_err_fp32_vs_fp64_per_root_median"] = [float(v) for v in np.median(err, axis=0)]
    return out


def pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q))


def serve_batched(reqs, opt, args, label, policy):
    """Serve one request stream with continuous batching under one admission
    policy; returns (results, stats)."""
    import threading
    import sbd_amd
    sv = sbd_amd.Server(args.serve_mem_gb << 30, max_active=args.serve_slots, policy=policy)
    ids = [None] * len(reqs)
    base = sv.now()

    def feeder():
        for q, r in enumerate(reqs):
            dt = base + r["t"] - sv.now()
            if dt > 0:
                time.sleep(dt)
            ids[q] = sv.submit(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"],
                               sbd=opt)

    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    res = []
    for q, i in enumerate(ids):
        res.append(sv.wait(i))
        if (q + 1) % 16 == 0:
            log(f"serving {label}: batched ({policy}), {q + 1} solves done")
    st = sv.stats()
    sv.close()
    lat = [r["t_done"] - r["t_submit"] for r in res]
    span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
    log(f"serving {label}: batched ({policy}) p50 {pct(lat, 50):.3f}s p99 {pct(lat, 99):.3f}s "
        f"{len(res) / span:.2f} solves/s peak active {st['peak_active']}")
    return res, {"solves_per_s": round(len(res) / span, 3), "p50_s": round(pct(lat, 50), 4),
                 "p99_s": round(pct(lat, 99), 4), "makespan_s": round(span, 3),
                 "peak_active": st["peak_active"], "rounds": st["rounds"],
                 "converged": sum(r["converged"] for r in res)}


def serve_stream(reqs, opt, args, label, policies):
    """Serve one request stream with continuous batching (one run per admission
    policy), then the same stream one request at a time with the plain solver
    (arrival order, queueing included)."""
    import sbd_amd
    batched = {p: serve_batched(reqs, opt, args, label, p) for p in policies}
    seq_lat, seq_E = [], []
    t_start = time.time()
    for r in reqs:
        now = time.time() - t_start
        if now < r["t"]:
            time.sleep(r["t"] - now)
        m = sbd_amd.MultTPB(r["norb"], r["a"], r["b"], I0=r["I0"], h=r["h"], eri=r["eri"])
        E, _, rep = m.diag(opt)
        m.close()
        seq_E.append(E)
        seq_lat.append(time.time() - t_start - r["t"])
        if len(seq_E) % 16 == 0:
            log(f"serving {label}: sequential, {len(seq_E)} solves done")
    seq_span = time.time() - t_start - reqs[0]["t"]
    log(f"serving {label}: sequential p50 {pct(seq_lat, 50):.3f}s p99 {pct(seq_lat, 99):.3f}s "
        f"{len(reqs) / seq_span:.2f} solves/s")
    seq = {"solves_per_s": round(len(reqs) / seq_span, 3), "p50_s": round(pct(seq_lat, 50), 4),
           "p99_s": round(pct(seq_lat, 99), 4), "makespan_s": round(seq_span, 3)}
    out = {"requests": len(reqs), "sequential_plain": seq}
    for p, (res, b) in batched.items():
        same = sum(1 for a, e in zip(res, seq_E) if a["energy"] == e)
        b["throughput_ratio"] = round(b["solves_per_s"] / seq["solves_per_s"], 3)
        b["p50_ratio_sequential_over_batched"] = round(seq["p50_s"] / max(b["p50_s"], 1e-9), 3)
        b["energies_bit_identical"] = f"{same}/{len(reqs)}"
        out[f"batched_{p}"] = b
    return out


def serving(args):
    """Poisson streams of independent solves (norb 16-40, n 4-10, 200-8000 strings
    per spin, per-request integrals), served with continuous batching, against the
    same stream solved one request at a time with the plain solver.  Two loads:
    "overload" (arrivals far faster than the GPU completes solves: throughput)
    and "steady" (about half the sequential service rate: latency, where a
    small solve that arrives behind a large one waits for it when solves run
    one at a time).  Batched runs use the server's admission policies "fifo"
    (arrival order) and "sjf" (smallest footprint first, with aging); the
    sequential baseline takes requests in arrival order."""
    import sbd_amd
    from sbd_amd import synth
    opt = sbd_amd.SBD(max_it=args.serve_max_it, max_nb=10, eps=args.serve_eps, init=0)
    out = {"options": {"max_it": args.serve_max_it, "max_nb": 10, "eps": args.serve_eps,
                       "init": 0}}
    pol = tuple(args.serve_policies.split(","))
    for label, count, rate, seed, policies in (
            ("overload", args.serve_count, args.serve_rate, 4004, pol),
            ("steady", args.serve_steady_count, args.serve_steady_rate, 4005, pol)):
        if count <= 0:
            continue
        t0 = time.time()
        reqs = synth.serving_requests(count, seed, rate=rate)
        log(f"serving {label}: {len(reqs)} requests at {rate}/s generated in "
            f"{time.time() - t0:.1f}s, N total {sum(len(r['a']) * len(r['b']) for r in reqs):.3e}")
        out[label] = serve_stream(reqs, opt, args, label, policies)
        out[label]["arrival_rate_per_s"] = rate
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="sci36")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-davidson", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--dav-restarts", type=int, default=60,
                    help="Davidson max_it (restarts of max_nb=10 steps each)")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-rdm", action="store_true")
    ap.add_argument("--serve-count", type=int, default=256,
                    help="requests in the overload stream")
    ap.add_argument("--serve-rate", type=float, default=40.0,
                    help="Poisson arrivals per second of the overload stream")
    ap.add_argument("--serve-steady-count", type=int, default=0,
                    help="requests in the steady (sub-saturation) stream; 0 skips it")
    ap.add_argument("--serve-steady-rate", type=float, default=0.5,
                    help="Poisson arrivals per second of the steady stream")
    ap.add_argument("--serve-slots", type=int, default=32)
    ap.add_argument("--serve-mem-gb", type=int, default=160)
    ap.add_argument("--serve-eps", type=float, default=1e-6)
    ap.add_argument("--serve-max-it", type=int, default=20)
    ap.add_argument("--serve-policies", default="sjf",
                    help="admission policies of the batched runs, comma-separated (fifo, sjf)")
    ap.add_argument("--no-dense", action="store_true")
    ap.add_argument("--cpu-davidson-s", type=float, default=60.0,
                    help="longest host Davidson (oracle, FCI norb 12) the dense section runs")
    ap.add_argument("--no-molecules", action="store_true",
                    help="dense section without the reference's molecule inputs")
    ap.add_argument("--large-molecules", nargs="?", const="h2o_1em7,n2_fci", default="",
                    help="dense section also on these of the reference's larger inputs, "
                         "comma-separated: h2o_1em7 (6.28e8 determinants), n2_fci (1.01e9), "
                         "h2o_1em8 (1.52e9; about 11 minutes); no list: the first two")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--prec-max-it", type=int, default=80,
                    help="Davidson restarts of the precision section (subspace 10, or what fits)")
    ap.add_argument("--prec-eps", type=float, default=1e-6,
                    help="residual of the precision section's solves")
    ap.add_argument("--no-multiroot", action="store_true")
    ap.add_argument("--mr-count", type=int, default=32)
    ap.add_argument("--mr-scale", type=float, default=0.1,
                    help="string-count scale of the multi-root requests (1.0 = 30000 x 25000)")
    ap.add_argument("--mr-max-it", type=int, default=60)
    ap.add_argument("--sigma-kernel", default="auto",
                    choices=["auto", "rows", "elements", "rows_inline", "split"],
                    help="sigma kernel of the timed step (all give the same bits)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="run only a plain 2 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box here has one GPU, so --gpus must be 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import sbd_amd
    from sbd_amd import synth

    torch.cuda.set_device(0)
    if args.copy_calib:
        n = (2 << 30) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n}))
        return

    t0 = time.time()
    w = synth.make_workload(args.workload, args.scale)
    log(f"workload {args.workload} x{args.scale}: na={len(w['a'])} nb={len(w['b'])} "
        f"generated in {time.time() - t0:.1f}s")
    t0 = time.time()
    m = sbd_amd.MultTPB(w["norb"], w["a"], w["b"], I0=w["I0"], h=w["h"], eri=w["eri"])
    torch.cuda.synchronize()
    t_tables = time.time() - t0
    counts = m.table_counts()
    N = m.N
    log(f"tables {counts.tolist()} in {t_tables:.2f}s; N={N:.3e}")
    m.set_sigma_kernel(args.sigma_kernel)
    t0 = time.time()
    hii = m.hdiag()
    torch.cuda.synchronize()
    t_hdiag = time.time() - t0
    g = torch.Generator(device="cuda")
    g.manual_seed(1234)
    T = torch.randn(N, dtype=torch.float64, device="cuda", generator=g)
    Wb = torch.empty_like(T)
    for _ in range(args.warmup):
        m.run(hii, T, Wb, accumulate=False)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.time()
    ev0.record(stream)
    for _ in range(args.steps):
        m.run(hii, T, Wb, accumulate=False)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.time() - wall0
    ms = ev0.elapsed_time(ev1) / args.steps
    gdet = N / (ms * 1e-3) / 1e9
    P_alpha = int(counts[0] + counts[1])
    s = 8
    B = s * (N + m.nb * P_alpha) + s * N
    achieved = B / (ms * 1e-3) / 1e9
    log(f"sigma {ms:.3f} ms/launch = {gdet:.2f} GDet/s, {achieved:.0f} GB/s algorithmic "
        f"({achieved / HBM_PEAK_GBS:.1%} of peak); wall {wall:.2f}s")

    traffic = None
    prof = os.path.join(ROOT, "profiles", f"sigma_{args.workload}_traffic.json")
    if os.path.exists(prof) and args.scale == 1.0:
        try:
            traffic = json.load(open(prof)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    out = {
        "metric": "sigma GDet/s + time-to-energy @1e-6 Ha; served solves/s with p50/p99 latency",
        "value": round(gdet, 4), "unit": "GDet/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded splitmix64 integrals and strings)",
        "config": {"workload": f"{args.workload}: {WORKLOAD_DESC.get(args.workload, '')}"
                               + (f" scaled x{args.scale}" if args.scale != 1.0 else ""),
                   "norb": w["norb"], "n_alpha_strings": m.na, "n_beta_strings": m.nb,
                   "determinants": N, "pairs_alpha": P_alpha,
                   "pairs_beta": int(counts[2] + counts[3]),
                   "step": "one FP64 sigma build over all determinants",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "algorithmic_bytes_per_launch": B,
                     "kernel": ("k_sigma_rows<double, true, true, 1, true>"
                                if args.sigma_kernel == "auto" and args.workload == "sci36"
                                else f"sigma kernel choice '{args.sigma_kernel}'")},
        "setup_s": {"tables": round(t_tables, 3), "hdiag": round(t_hdiag, 3)},
    }

    if not args.no_davidson:
        # each solve runs twice (the same iterates and energy both times); the
        # faster run is reported and both times are listed: a one-off slowdown
        # of a few seconds on the box then cannot decide the ratio
        def tte_of(hist, Eref):
            for E0, _, t, p in hist:
                if p == 64 and abs(E0 - Eref) <= 1e-6:
                    return t
            return None

        runs = {}
        for rep_i in range(2):
            for label, prec in (("fp64", 0), ("mixed", 1)):
                opt = sbd_amd.SBD(max_it=args.dav_restarts, max_nb=10, eps=1e-7, precision=prec)
                E, _, rep, hist = m.diag(opt, history=True)
                runs.setdefault(label, []).append(
                    {"energy": E, "steps": rep["steps"], "steps_fp32": rep["steps_fp32"],
                     "converged": rep["converged"], "residual": rep["residual"],
                     "t_davidson": rep["t_davidson"], "hist": hist})
                log(f"davidson {label} run {rep_i}: E={E:.10f} steps={rep['steps']} "
                    f"t={rep['t_davidson']:.2f}s conv={rep['converged']}")
        Eref = runs["fp64"][0]["energy"]

        def step_ms(hist):
            # wall time between consecutive Ritz values: min / median / max and
            # the step of the slowest (tells a uniformly slow run from a stall)
            import numpy as np
            t = np.asarray([row[2] for row in hist])
            if t.size < 2:
                return None
            dt = np.diff(t) * 1e3
            return {"min": round(float(dt.min()), 2), "median": round(float(np.median(dt)), 2),
                    "max": round(float(dt.max()), 2), "argmax": int(dt.argmax()) + 1}

        dav = {}
        for label, rs in runs.items():
            ttes = [tte_of(r["hist"], Eref) for r in rs]
            best = min(range(len(rs)), key=lambda i: (ttes[i] is None, ttes[i] or 0.0))
            r, tte = rs[best], ttes[best]
            dav[label] = {"time_to_energy_s": None if tte is None else round(tte, 4),
                          "time_to_energy_runs_s": [None if x is None else round(x, 4)
                                                    for x in ttes],
                          "step_ms_runs": [step_ms(x["hist"]) for x in rs],
                          "energy": r["energy"], "abs_err_vs_fp64": abs(r["energy"] - Eref),
                          "energies_equal_across_runs": len({x["energy"] for x in rs}) == 1,
                          "steps": r["steps"], "steps_fp32": r["steps_fp32"],
                          "converged": bool(r["converged"]), "residual": r["residual"],
                          "t_davidson_s": round(r["t_davidson"], 4)}
        t64, tmx = dav["fp64"]["time_to_energy_s"], dav["mixed"]["time_to_energy_s"]
        dav["speedup_mixed_vs_fp64"] = (round(t64 / tmx, 3) if t64 and tmx else None)
        dav["eps"] = 1e-7
        dav["runs"] = "each solve twice, the faster reported"
        out["davidson"] = dav

    if not args.no_cpu:
        log("cpu baseline (oracle) ...")
        m.run(hii, T, Wb, accumulate=False)
        torch.cuda.synchronize()
        out["cpu_baseline"] = cpu_baseline(w, m, T, Wb, hii, args.cpu_seconds)
        log(f"cpu baseline {out['cpu_baseline']['value']:.4f} GDet/s "
            f"(bit-exact rows: {out['cpu_baseline']['sample_bitexact_vs_gpu']})")
    else:
        out["cpu_baseline"] = None
    if not args.no_rdm:
        # full one- and two-particle RDMs (do_rdm != 0) of the normalised input
        # vector: wall time of one sbd_correlation call (device vector in, host
        # matrices out), median of three
        x = T / torch.linalg.norm(T)
        times = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            one, two = m.correlation(x)
            times.append(time.perf_counter() - t0)
        L = w["norb"]
        tr = [float(np.trace(one[s].reshape(L, L))) for s in range(2)]
        out["rdm"] = {"ms_per_call": round(1e3 * sorted(times)[1], 2), "calls": 3,
                      "trace_alpha": tr[0], "trace_beta": tr[1],
                      "outputs": "onebody 2 L^2 and twobody 4 L^4 doubles on the host"}
        log(f"rdm: {out['rdm']['ms_per_call']:.1f} ms per call; traces {tr[0]:.12f} {tr[1]:.12f}")
        del x, one, two
    m.close()
    del T, Wb, hii
    torch.cuda.empty_cache()
    for name, skip, fn in (("dense", args.no_dense, dense_section),
                           ("precision", args.no_precision, precision_section),
                           ("serving", args.no_serving, serving),
                           ("multiroot", args.no_multiroot, multiroot_section)):
        if skip:
            continue
        t0 = time.time()
        out[name] = fn(args)
        torch.cuda.empty_cache()
        log(f"section {name} done in {time.time() - t0:.1f}s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
