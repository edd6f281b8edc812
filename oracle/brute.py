"""TEST INFRASTRUCTURE ONLY: independent brute-force Hamiltonian for tiny systems.

Builds the dense CI matrix from second quantisation, with no code shared with
the Slater-Condon restatement:

    H = I0 + sum_{pq,s} h_pq a+_{ps} a_{qs}
           + 1/2 sum_{pqrs,st} (pq|rs) a+_{ps} a+_{rt} a_{st} a_{qs}

Spin orbitals are interleaved (k = 2p + s, s = 0 alpha, 1 beta), the
reference's full-determinant order (chemistry/basic/determinants.h:104-150).
A determinant is a+_{k1} a+_{k2} ... |0> with k1 < k2 < ...; a_k and a+_k
pick up (-1)^(number of occupied spin orbitals below k).  The basis is the
tensor product of the given alpha and beta strings, row index ia*nb + ib
(tpb/davidson.h:171-173).
"""
from __future__ import annotations

import numpy as np


def _annihilate(det: int, k: int):
    if not (det >> k) & 1:
        return 0, 0
    sgn = -1 if bin(det & ((1 << k) - 1)).count("1") % 2 else 1
    return sgn, det & ~(1 << k)


def _create(det: int, k: int):
    if (det >> k) & 1:
        return 0, 0
    sgn = -1 if bin(det & ((1 << k) - 1)).count("1") % 2 else 1
    return sgn, det | (1 << k)


def interleave(a: int, b: int, L: int) -> int:
    d = 0
    for p in range(L):
        if (a >> p) & 1:
            d |= 1 << (2 * p)
        if (b >> p) & 1:
            d |= 1 << (2 * p + 1)
    return d


def dense_hamiltonian(L, I0, h, eri, astr, bstr):
    """eri: dense (L,L,L,L) chemist-notation (pq|rs)."""
    dets = [interleave(int(a), int(b), L) for a in astr for b in bstr]
    index = {d: i for i, d in enumerate(dets)}
    n = len(dets)
    H = np.zeros((n, n))
    M = 2 * L
    for col, ket in enumerate(dets):
        H[col, col] += I0
        occ = [k for k in range(M) if (ket >> k) & 1]
        # one-body
        for q in occ:
            s1, d1 = _annihilate(ket, q)
            for p in range(q % 2, M, 2):
                s2, d2 = _create(d1, p)
                if s2 == 0 or d2 not in index:
                    continue
                H[index[d2], col] += s1 * s2 * h[p // 2, q // 2]
        # two-body: a+_p a+_r a_s a_q  (q, s annihilated first: a_q then a_s)
        for q in occ:
            sq, dq = _annihilate(ket, q)
            for s in occ:
                ss, ds = _annihilate(dq, s)
                if ss == 0:
                    continue
                for r in range(s % 2, M, 2):
                    sr, dr = _create(ds, r)
                    if sr == 0:
                        continue
                    for p in range(q % 2, M, 2):
                        sp, dp = _create(dr, p)
                        if sp == 0 or dp not in index:
                            continue
                        H[index[dp], col] += 0.5 * sq * ss * sr * sp * eri[p // 2, q // 2, r // 2, s // 2]
    return H


def _apply(det: int, seq):
    """Apply operators in order (first element acts first): ('a', k) or ('c', k)."""
    sgn = 1
    for kind, k in seq:
        s, det = (_annihilate if kind == "a" else _create)(det, k)
        if s == 0:
            return 0, 0
        sgn *= s
    return sgn, det


def _swap_class(X: int, Y: int) -> bool:
    """An alpha x beta double excitation between X and Y whose lowest X-only
    spin orbital has the spin of the highest Y-only one (symmetric in X, Y)."""
    xo, yo = X & ~Y, Y & ~X
    if bin(xo).count("1") != 2 or bin(yo).count("1") != 2:
        return False
    lo_x = (xo & -xo).bit_length() - 1
    hi_y = yo.bit_length() - 1
    spins_x = {k % 2 for k in range(lo_x, xo.bit_length()) if (xo >> k) & 1}
    return spins_x == {0, 1} and lo_x % 2 == hi_y % 2


def rdm(L, astr, bstr, W):
    """One- and two-particle RDMs of the real vector W on the tensor-product
    basis, from second quantisation:
        one[s, q, p]               = <a+_{ps} a_{qs}>
        two[s1 + 2 s2, t, r, q, p] = <a+_{p s1} a+_{q s2} a_{t s2} a_{r s1}>
    (numpy index order reversed, so that the flat C-order index is
    p + L q + L^2 r + L^3 t).  Also returns the part of `two` contributed by
    alpha x beta double excitations of the `_swap_class` kind."""
    dets = [interleave(int(a), int(b), L) for a in astr for b in bstr]
    index = {d: i for i, d in enumerate(dets)}
    one = np.zeros((2, L, L))
    two = np.zeros((4, L, L, L, L))
    swap = np.zeros((4, L, L, L, L))
    for J, ket in enumerate(dets):
        if W[J] == 0.0:
            continue
        occ = [k for k in range(2 * L) if (ket >> k) & 1]
        for q in occ:
            s = q % 2
            for p in range(s, 2 * L, 2):
                sg, d = _apply(ket, [("a", q), ("c", p)])
                if sg and d in index:
                    one[s, q // 2, p // 2] += W[index[d]] * W[J] * sg
        for r in occ:
            for t in occ:
                if t == r:
                    continue
                for q in range(t % 2, 2 * L, 2):
                    for p in range(r % 2, 2 * L, 2):
                        sg, d = _apply(ket, [("a", r), ("a", t), ("c", q), ("c", p)])
                        if not sg or d not in index:
                            continue
                        blk = r % 2 + 2 * (t % 2)
                        v = W[index[d]] * W[J] * sg
                        two[blk, t // 2, r // 2, q // 2, p // 2] += v
                        if _swap_class(d, ket):
                            swap[blk, t // 2, r // 2, q // 2, p // 2] += v
    return one, two, swap
