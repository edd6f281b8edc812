"""TEST INFRASTRUCTURE ONLY: Python side of the CPU oracle.

Restates the reference's input handling and wraps the C restatement
(oracle/sbd_oracle.c -> liboracle.so) of the TPB path.  Used by tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg as the checker; the
product path never imports it.

* load_fcidump   : framework/fcidump.h:29-95
* from_string    : framework/bit_manipulation.h:219-235 (bit_length = 64)
* load_alpha_dets: chemistry/basic/makedeterminants.h:167-191 + sort_bitarray
                   (bit_manipulation.h:421-449; ascending value, unique)
* setup_integrals: chemistry/basic/makeintegrals.h:124-174 (in C)
* System         : tables (tpb/helper.h:185-342), hii (tpb/qcham.h:290-340),
                   sigma (tpb/mult.h:625-881), Davidson (tpb/davidson.h:707-970),
                   Lanczos (tpb/lanczos.h:162-327), weights / occupations /
                   carry-over (tpb/rdmat.h, tpb/occupation.h), RDMs
                   (tpb/correlation.h:379-643)
"""
from __future__ import annotations

import ctypes as C
import gzip
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp, lp, ip, up = (C.POINTER(C.c_double), C.POINTER(C.c_long), C.POINTER(C.c_int),
                          C.POINTER(C.c_uint64))
        L.orc_setup_integrals.argtypes = [C.c_int, C.c_long, dp, ip, ip, ip, ip, dp, dp, dp, dp, dp]
        L.orc_sys_create.restype = C.c_void_p
        L.orc_sys_create.argtypes = [C.c_int, up, C.c_long, up, C.c_long, C.c_double, dp, dp, dp, dp]
        L.orc_sys_free.argtypes = [C.c_void_p]
        L.orc_sys_counts.argtypes = [C.c_void_p, lp]
        L.orc_sys_table.argtypes = [C.c_void_p, C.c_int, lp, lp, ip]
        L.orc_hdiag.argtypes = [C.c_void_p, dp]
        L.orc_sigma.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_inner.restype = C.c_double
        L.orc_inner.argtypes = [dp, dp, C.c_long]
        L.orc_normalize.restype = C.c_double
        L.orc_normalize.argtypes = [dp, C.c_long]
        L.orc_syev.argtypes = [C.c_int, dp, C.c_int, dp]
        L.orc_davidson.restype = C.c_int
        L.orc_davidson.argtypes = [C.c_void_p, dp, dp, C.c_int, C.c_int, C.c_double, dp]
        L.orc_lanczos.argtypes = [C.c_void_p, dp, dp, C.c_int, C.c_int, C.c_double, dp]
        L.orc_weights.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_occupation.argtypes = [C.c_void_p, dp, dp]
        L.orc_carryover.restype = C.c_long
        L.orc_carryover.argtypes = [C.c_void_p, dp, C.c_int, C.c_double, C.c_double, lp, dp]
        L.orc_correlation.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_hdiag_rows.argtypes = [C.c_void_p, dp, C.c_long, C.c_long]
        L.orc_sigma_rows.argtypes = [C.c_void_p, dp, dp, dp, C.c_long, C.c_long]
        L.orc_det_from_ab.argtypes = [up, up, C.c_long, C.c_int, up]
        L.orc_zero_excite.restype = C.c_double
        L.orc_zero_excite.argtypes = [up, C.c_int, C.c_double, dp, dp, dp]
        L.orc_one_excite.restype = C.c_double
        L.orc_one_excite.argtypes = [up, C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_two_excite.restype = C.c_double
        L.orc_two_excite.argtypes = [up, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.orc_parity.restype = C.c_double
        L.orc_parity.argtypes = [up, C.c_int, C.c_int]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _open(path):
    return gzip.open(path, "rt") if path.endswith(".gz") else open(path)


def load_fcidump(path):
    """fcidump.h:29-95: header key/values (comma separated between &FCI and &END),
    then `value i j k l` lines in file order."""
    header_text = ""
    vals, idx = [], []
    in_header = True
    with _open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line or line[0] == "!":
                continue
            if in_header:
                if "&END" in line:
                    in_header = False
                    continue
                p = line.find("&FCI")
                if p >= 0:
                    line = line[:p] + line[p + 4:]
                header_text += line
            else:
                t = line.split()
                if len(t) < 5:
                    continue
                vals.append(float(t[0]))
                idx.append((int(t[1]), int(t[2]), int(t[3]), int(t[4])))
    header = {}
    for tok in header_text.split(","):
        if "=" in tok:
            k, v = tok.split("=", 1)
            header[k.strip(" \t")] = v.strip(" \t")
    return header, np.array(vals, dtype=np.float64), np.array(idx, dtype=np.int32).reshape(-1, 4)


def setup_integrals(header, vals, idx):
    """makeintegrals.h:124-174 (RHF).  Returns (L, nelec, I0, I1, I2, Dm, Xm)."""
    L = int(header["NORB"])
    nelec = int(header.get("NELEC", "0"))
    B = L * (L + 1) // 2
    I0 = C.c_double(0.0)
    I1 = np.zeros(4 * L * L)
    I2 = np.zeros(B * (B + 1) // 2)
    Dm = np.zeros(4 * L * L)
    Xm = np.zeros(4 * L * L)
    cols = [np.ascontiguousarray(idx[:, k]) for k in range(4)]
    lib().orc_setup_integrals(L, len(vals), _p(vals, C.c_double), *[_p(c, C.c_int) for c in cols],
                              C.byref(I0), _p(I1, C.c_double), _p(I2, C.c_double),
                              _p(Dm, C.c_double), _p(Xm, C.c_double))
    return L, nelec, I0.value, I1, I2, Dm, Xm


def integrals_from_h_eri(L, I0, h, eri_packed):
    """Build I1/Dm/Xm from (h, packed ERI) through the same setup routine."""
    B = L * (L + 1) // 2
    vals, idx = [], []
    pairs = [(i, j) for i in range(L) for j in range(i + 1)]
    nz = np.nonzero(eri_packed)[0]
    a = np.floor((np.sqrt(8 * nz + 1) - 1) / 2).astype(np.int64)
    b = nz - a * (a + 1) // 2
    for ai, bi, v in zip(a.tolist(), b.tolist(), eri_packed[nz].tolist()):
        (i, j), (k, l) = pairs[ai], pairs[bi]
        vals.append(v)
        idx.append((i + 1, j + 1, k + 1, l + 1))
    for i in range(L):
        for j in range(i + 1):
            if h[i, j] != 0.0:
                vals.append(float(h[i, j]))
                idx.append((i + 1, j + 1, 0, 0))
    vals.append(float(I0))
    idx.append((0, 0, 0, 0))
    hdr = {"NORB": str(L), "NELEC": "0"}
    assert len(eri_packed) == B * (B + 1) // 2
    return setup_integrals(hdr, np.array(vals), np.array(idx, dtype=np.int32))


def from_string(s: str, L: int) -> int:
    """bit_manipulation.h:219-235: character s[L-1-i] is orbital i."""
    v = 0
    for i in range(L):
        if s[L - 1 - i] == "1":
            v |= 1 << i
    return v


def load_alpha_dets(path, L):
    """makedeterminants.h:167-191 (txt) + sort_bitarray: ascending, unique."""
    with _open(path) as f:
        vals = [from_string(line.rstrip("\n"), L) for line in f]
    return np.array(sorted(set(vals)), dtype=np.uint64)


class System:
    """CPU oracle of one TPB system (one rank)."""

    def __init__(self, L, a, b, I0, I1, I2, Dm, Xm):
        self.L = L
        self.a = np.ascontiguousarray(a, dtype=np.uint64)
        self.b = np.ascontiguousarray(b, dtype=np.uint64)
        self.na, self.nb = len(self.a), len(self.b)
        self.N = self.na * self.nb
        self.I0, self.I1, self.I2, self.Dm, self.Xm = I0, I1, I2, Dm, Xm
        self._h = lib().orc_sys_create(L, _p(self.a, C.c_uint64), self.na, _p(self.b, C.c_uint64),
                                       self.nb, I0, _p(I1, C.c_double), _p(I2, C.c_double),
                                       _p(Dm, C.c_double), _p(Xm, C.c_double))

    @classmethod
    def from_workload(cls, w):
        L, _, I0, I1, I2, Dm, Xm = integrals_from_h_eri(w["norb"], w["I0"], w["h"], w["eri"])
        return cls(L, w["a"], w["b"], I0, I1, I2, Dm, Xm)

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().orc_sys_free(self._h)
            except (TypeError, AttributeError):  # interpreter shutdown
                pass
            self._h = None

    def counts(self):
        out = np.zeros(4, dtype=np.int64)
        lib().orc_sys_counts(self._h, _p(out, C.c_long))
        return out

    def table(self, which):
        """which: 0 alpha singles, 1 alpha doubles, 2 beta singles, 3 beta doubles.
        Returns (offsets, kets, CrAn) in the reference's layout."""
        n = self.na if which < 2 else self.nb
        cnt = self.counts()[which]
        w = 2 if which % 2 == 0 else 4
        off = np.zeros(n + 1, dtype=np.int64)
        ket = np.zeros(max(cnt, 1), dtype=np.int64)
        cr = np.zeros(max(w * cnt, 1), dtype=np.int32)
        lib().orc_sys_table(self._h, which, _p(off, C.c_long), _p(ket, C.c_long), _p(cr, C.c_int))
        return off, ket[:cnt], cr[:w * cnt].reshape(cnt, w)

    def hdiag(self):
        out = np.zeros(self.N)
        lib().orc_hdiag(self._h, _p(out, C.c_double))
        return out

    def sigma(self, hii, T, Wb=None):
        """Wb += H T (accumulating, like MultBase::run)."""
        T = np.ascontiguousarray(T, dtype=np.float64)
        Wb = np.zeros(self.N) if Wb is None else np.ascontiguousarray(Wb, dtype=np.float64).copy()
        lib().orc_sigma(self._h, _p(hii, C.c_double), _p(T, C.c_double), _p(Wb, C.c_double))
        return Wb

    def hdiag_rows(self, hii, ia0, ia1):
        """Fill rows [ia0, ia1) of a full-size hii (in place)."""
        lib().orc_hdiag_rows(self._h, _p(hii, C.c_double), ia0, ia1)

    def sigma_rows(self, hii, T, Wb, ia0, ia1):
        """Wb[rows ia0..ia1) += (H T)[rows] on full-size vectors (in place)."""
        lib().orc_sigma_rows(self._h, _p(hii, C.c_double), _p(T, C.c_double),
                             _p(Wb, C.c_double), ia0, ia1)

    def davidson(self, hii, W, max_it, nb, eps):
        W = np.ascontiguousarray(W, dtype=np.float64).copy()
        hist = np.zeros(2 * max_it * nb)
        steps = lib().orc_davidson(self._h, _p(hii, C.c_double), _p(W, C.c_double), max_it, nb,
                                   eps, _p(hist, C.c_double))
        return W, hist[:2 * steps].reshape(steps, 2)

    def lanczos(self, hii, W, max_it, nb, eps):
        """orc_lanczos (lanczos.h:162-327): returns W and the per-step history
        rows (E0, alpha, beta)."""
        W = np.ascontiguousarray(W, dtype=np.float64).copy()
        hist = np.zeros(3 * max_it * nb)
        steps = lib().orc_lanczos(self._h, _p(hii, C.c_double), _p(W, C.c_double), max_it, nb,
                                  eps, _p(hist, C.c_double))
        return W, hist[:3 * steps].reshape(steps, 3)

    def weights(self, W):
        """rdmat.h:11-58, 136-182: (alpha string weights, beta string weights)."""
        W = np.ascontiguousarray(W, dtype=np.float64)
        Da, Db = np.zeros(self.na), np.zeros(self.nb)
        lib().orc_weights(self._h, _p(W, C.c_double), _p(Da, C.c_double), _p(Db, C.c_double))
        return Da, Db

    def occupation(self, W):
        """occupation.h:10-95: density[2o] alpha, density[2o+1] beta."""
        W = np.ascontiguousarray(W, dtype=np.float64)
        out = np.zeros(2 * self.L)
        lib().orc_occupation(self._h, _p(W, C.c_double), _p(out, C.c_double))
        return out

    def carryover(self, W, spin, threshold, ratio=0.0):
        """rdmat.h:60-134: indices of the kept strings and the truncated weight."""
        W = np.ascontiguousarray(W, dtype=np.float64)
        out = np.zeros(self.nb if spin else self.na, dtype=np.int64)
        tr = C.c_double(0.0)
        m = lib().orc_carryover(self._h, _p(W, C.c_double), spin, threshold, ratio,
                                out.ctypes.data_as(C.POINTER(C.c_long)), C.byref(tr))
        return out[:m], tr.value

    def correlation(self, W):
        """correlation.h:379-643 (do_rdm != 0): (onebody (2, L, L), twobody
        (4, L, L, L, L)); element [s][o1 + L*o2] is onebody[s, o2, o1] and
        [s1 + 2*s2][o1 + L*o2 + L^2*o3 + L^3*o4] is twobody[s1 + 2*s2, o4, o3, o2, o1]."""
        W = np.ascontiguousarray(W, dtype=np.float64)
        L = self.L
        one, two = np.zeros(2 * L * L), np.zeros(4 * L ** 4)
        lib().orc_correlation(self._h, _p(W, C.c_double), _p(one, C.c_double),
                              _p(two, C.c_double))
        return one.reshape(2, L, L), two.reshape(4, L, L, L, L)

    def energy(self, hii, W):
        """sbdiag.h:488-526: E = <W, H W> with a fresh sigma."""
        HW = self.sigma(hii, W)
        return inner(W, HW)


def inner(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().orc_inner(_p(x, C.c_double), _p(y, C.c_double), len(x))


def normalize(x):
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    n = lib().orc_normalize(_p(x, C.c_double), len(x))
    return x, n


def syev(A):
    A = np.asfortranarray(np.array(A, dtype=np.float64))
    n = A.shape[0]
    buf = np.ascontiguousarray(A.T).copy()   # column-major buffer
    w = np.zeros(n)
    lib().orc_syev(n, _p(buf, C.c_double), n, _p(w, C.c_double))
    return w, buf.reshape(n, n).T.copy()
