"""Seeded synthetic inputs for the tensor-product-basis (TPB) workloads.

One portable generator is defined here and used by the benchmarks and tests:

* RNG: splitmix64 (Steele, Lea, Flood 2014) streams; uniforms take the top 53
  bits; normals come from the Box-Muller transform, both outputs of each pair
  used in order.
* one-body h: symmetric, off-diagonal N(0, 0.5), diagonal linspace(-2, 0, L).
* two-body (pq|rs) = sum_k L^k_pq L^k_rs with K symmetric factors whose upper
  triangles are N(0, 0.3): 8-fold symmetric and positive semidefinite.
* strings: the Hartree-Fock string (orbitals 0..n-1) plus uniform random
  weight-n strings drawn by a partial Fisher-Yates shuffle, deduplicated and
  sorted ascending (the reference's sort order, bit_manipulation.h:153-171,
  421-449).

The integral layouts produced are the reference's RHF layouts
(chemistry/basic/makeintegrals.h:124-174): packed 8-fold ERI of size B(B+1)/2
with B = L(L+1)/2, index ij = max(i,j)(max(i,j)+1)/2 + min(i,j).
"""
from __future__ import annotations

import math
from itertools import combinations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """splitmix64 stream; ``draw(n)`` returns the next n 64-bit outputs."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def draw(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * _GAMMA
            self.state = self.state + np.uint64(n) * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int) -> np.ndarray:
        return (self.draw(n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u = self.uniform(2 * m)
        u1, u2 = 1.0 - u[0::2], u[1::2]          # u1 in (0, 1]
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.empty(2 * m)
        z[0::2] = r * np.cos(2.0 * np.pi * u2)
        z[1::2] = r * np.sin(2.0 * np.pi * u2)
        return z[:n]


def elem_hash(seed: int, idx: np.ndarray) -> np.ndarray:
    """Per-element hash of the reference's Randomize (framework/dm_vector.h:102-113):
    a linear mix of (seed, index) followed by the splitmix64 output finaliser."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = np.uint64(seed) * np.uint64(6364136223846793005) + idx * np.uint64(2862933555777941757)
        x ^= x >> np.uint64(30)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(27)
        x *= np.uint64(0x94D049BB133111EB)
        x ^= x >> np.uint64(31)
    return x


def randomize(N: int, seed: int) -> np.ndarray:
    """The init = 1 start vector (BasisInitVector -> Randomize, davidson.h:178-181,
    dm_vector.h:114-191) at one rank and one thread: element i = the top 53 bits
    of elem_hash(seed, i) mapped to [-1, 1), scaled by 1/sqrt of the squared norm
    summed with Kahan compensation in index order (seed > 0)."""
    h = elem_hash(seed, np.arange(N, dtype=np.uint64))
    w = (h >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0
    s = e = 0.0
    for x in (w * w).tolist():
        v = x - e
        t = s + v
        e = (t - s) - v
        s = t
    return w * (1.0 / math.sqrt(s))


def random_init_vector(N: int, seed: int) -> np.ndarray:
    """A seeded random test vector: splitmix64(seed) uniforms mapped to (-1, 1),
    normalised (the init = 1 start vector is `randomize`)."""
    w = 2.0 * SplitMix64(seed).uniform(N) - 1.0
    return w / np.sqrt(np.dot(w, w))


def eri_index(i, j, k, l):
    ij = max(i, j) * (max(i, j) + 1) // 2 + min(i, j)
    kl = max(k, l) * (max(k, l) + 1) // 2 + min(k, l)
    a, b = max(ij, kl), min(ij, kl)
    return a * (a + 1) // 2 + b


def packed_size(L: int) -> int:
    B = L * (L + 1) // 2
    return B * (B + 1) // 2


def pack_eri(eri: np.ndarray) -> np.ndarray:
    """Dense (L,L,L,L) 8-fold symmetric ERI -> packed vector (reference layout)."""
    L = eri.shape[0]
    iu, ju = np.tril_indices(L)            # i >= j, row-major order = pair index ij
    pairs = eri[iu, ju][:, iu, ju]         # (B, B) over pair indices
    a, b = np.tril_indices(pairs.shape[0])
    return np.ascontiguousarray(pairs[a, b])


def unpack_eri(packed: np.ndarray, L: int) -> np.ndarray:
    B = L * (L + 1) // 2
    P = np.zeros((B, B))
    a, b = np.tril_indices(B)
    P[a, b] = packed
    P[b, a] = packed
    iu, ju = np.tril_indices(L)
    pid = np.zeros((L, L), dtype=np.int64)
    pid[iu, ju] = np.arange(B)
    pid[ju, iu] = np.arange(B)
    return P[pid[:, :, None, None], pid[None, None, :, :]]


def make_integrals(L: int, seed: int, rank: int | None = None):
    """Synthetic integrals: returns (I0, h (L,L), packed ERI)."""
    if rank is None:
        rank = 2 * L
    rng = SplitMix64(seed)
    h = np.zeros((L, L))
    iu = np.triu_indices(L, 1)
    off = 0.5 * rng.normal(len(iu[0]))
    h[iu] = off
    h[(iu[1], iu[0])] = off
    h[np.arange(L), np.arange(L)] = np.linspace(-2.0, 0.0, L)
    tu = np.triu_indices(L)
    Lk = np.zeros((rank, L, L))
    for k in range(rank):
        v = 0.3 * rng.normal(len(tu[0]))
        Lk[k][tu] = v
        Lk[k][(tu[1], tu[0])] = v
    Lm = Lk.reshape(rank, L * L)
    eri = (Lm.T @ Lm).reshape(L, L, L, L)
    return 0.0, h, pack_eri(eri)


def all_strings(L: int, n: int) -> np.ndarray:
    """All C(L,n) weight-n strings, ascending (FCI)."""
    out = [sum(1 << p for p in c) for c in combinations(range(L), n)]
    return np.array(sorted(out), dtype=np.uint64)


def random_strings(L: int, n: int, count: int, seed: int) -> np.ndarray:
    """HF string + uniform random weight-n strings, deduplicated, sorted ascending."""
    total = math.comb(L, n)
    if count > total:
        raise ValueError(f"count {count} exceeds C({L},{n}) = {total}")
    if count == total:
        return all_strings(L, n)
    rng = SplitMix64(seed)
    hf = (1 << n) - 1
    seen = {hf}
    while len(seen) < count:
        m = max(1024, 2 * (count - len(seen)))
        r = rng.draw(m * n)
        perm = np.tile(np.arange(L, dtype=np.int64), (m, 1))
        rows = np.arange(m)
        for i in range(n):
            j = i + (r[i::n][:m] % np.uint64(L - i)).astype(np.int64)
            a = perm[rows, i].copy()
            perm[rows, i] = perm[rows, j]
            perm[rows, j] = a
        bits = (np.uint64(1) << perm[:, :n].astype(np.uint64))
        vals = np.bitwise_or.reduce(bits, axis=1)
        for v in vals.tolist():
            if len(seen) >= count:
                break
            seen.add(int(v))
    return np.array(sorted(seen), dtype=np.uint64)


# Workloads: norb, n_alpha, n_beta, alpha strings, beta strings, ERI rank, seed.
WORKLOADS = {
    "fci12": dict(norb=12, na=4, nb=4, ca=495, cb=495, rank=24, seed=1001),
    "sci36": dict(norb=36, na=8, nb=8, ca=20000, cb=20000, rank=72, seed=2002),
    "sci58": dict(norb=58, na=10, nb=10, ca=60000, cb=60000, rank=116, seed=3003),
    "sci48open": dict(norb=48, na=12, nb=10, ca=30000, cb=25000, rank=96, seed=5005),
    # small, densely connected selected bases for Davidson parity (in the sparse
    # workloads above most determinants are nearly isolated at small scale)
    "sci16": dict(norb=16, na=4, nb=4, ca=400, cb=400, rank=32, seed=6006),
    "sci16open": dict(norb=16, na=5, nb=3, ca=400, cb=300, rank=32, seed=7007),
    # the sci58 integrals and sampled strings (1 % of them) plus every single
    # excitation of the Hartree-Fock string: in the sampled sci58 basis the HF
    # determinant has no coupled partner, so this is the connected norb-58 basis
    # used for precision parity
    "sci58hf": dict(norb=58, na=10, nb=10, ca=600, cb=600, rank=116, seed=3003, hf_singles=True),
    # the same for the open-shell sci48open basis (1 % of its sampled strings
    # plus the HF singles): the connected basis for multi-root parity
    "sci48openhf": dict(norb=48, na=12, nb=10, ca=300, cb=250, rank=96, seed=5005,
                        hf_singles=True),
}


def make_workload(name: str, scale: float = 1.0):
    """Return dict(norb, I0, h, eri, a, b) for a named workload.

    ``scale`` < 1 shrinks the string counts (same generator and seeds) for
    quick parity runs."""
    w = WORKLOADS[name]
    L = w["norb"]
    I0, h, eri = make_integrals(L, w["seed"], w["rank"])
    ca = max(2, int(round(w["ca"] * scale)))
    cb = max(2, int(round(w["cb"] * scale)))
    a = random_strings(L, w["na"], ca, w["seed"] + 1)
    b = random_strings(L, w["nb"], cb, w["seed"] + 2)
    if w.get("hf_singles"):
        a, b = (np.array(sorted(set(x.tolist()) | set(hf_singles(L, n).tolist())), dtype=np.uint64)
                for x, n in ((a, w["na"]), (b, w["nb"])))
    return dict(norb=L, I0=I0, h=h, eri=eri, a=a, b=b)


def hf_singles(L: int, n: int) -> np.ndarray:
    """The single excitations of the Hartree-Fock string (orbitals 0..n-1)."""
    hf = (1 << n) - 1
    return np.array(sorted(hf ^ (1 << i) ^ (1 << a) for i in range(n) for a in range(n, L)),
                    dtype=np.uint64)


def write_fcidump(path: str, L: int, nelec: int, I0: float, h: np.ndarray, eri: np.ndarray,
                  tol: float = 0.0):
    """Write an FCIDUMP (chemist notation, 1-based, one entry per 8-fold class)."""
    with open(path, "w") as f:
        f.write(f" &FCI NORB={L:4d},NELEC={nelec},MS2=0,\n")
        f.write("  ORBSYM=" + "1," * L + "\n  ISYM=1,\n &END\n")
        B = L * (L + 1) // 2
        pairs = [(i, j) for i in range(L) for j in range(i + 1)]
        for a in range(B):
            i, j = pairs[a]
            for b in range(a + 1):
                k, l = pairs[b]
                v = eri[a * (a + 1) // 2 + b]
                if abs(v) > tol:
                    f.write(f"{v!r} {i + 1} {j + 1} {k + 1} {l + 1}\n")
        for i in range(L):
            for j in range(i + 1):
                if abs(h[i, j]) > tol:
                    f.write(f"{h[i, j]!r} {i + 1} {j + 1} 0 0\n")
        f.write(f"{I0!r} 0 0 0 0\n")


def serving_requests(count: int, seed: int, norb=(16, 40), nelec=(4, 10), strings=(200, 8000),
                     rate: float = 0.0):
    """A stream of independent solve requests: norb uniform in [norb], n_alpha =
    n_beta = n uniform in [nelec], alpha and beta string counts log-uniform in
    [strings] (capped at C(norb, n)), random weight-n strings plus HF, integrals of
    rank 2 norb from a per-request seed.  With rate > 0 each request carries a
    Poisson arrival time (exponential gaps of mean 1/rate seconds) in "t"."""
    rng = SplitMix64(seed)
    out, t = [], 0.0
    for q in range(count):
        r = rng.draw(6)
        u = (r >> np.uint64(11)).astype(np.float64) / 9007199254740992.0
        L = norb[0] + int(r[0] % np.uint64(norb[1] - norb[0] + 1))
        n = nelec[0] + int(r[1] % np.uint64(nelec[1] - nelec[0] + 1))
        n = min(n, L - 1)
        cap = math.comb(L, n)
        lo, hi = math.log(strings[0]), math.log(strings[1])
        ca = min(cap, int(round(math.exp(lo + (hi - lo) * u[2]))))
        cb = min(cap, int(round(math.exp(lo + (hi - lo) * u[3]))))
        s0 = int(r[4] >> np.uint64(1))
        if rate > 0:
            t += -math.log(max(u[5], 1e-300)) / rate
        I0, h, eri = make_integrals(L, s0, 2 * L)
        out.append(dict(id=q, norb=L, n=n, I0=I0, h=h, eri=eri, t=t,
                        a=random_strings(L, n, ca, s0 + 1), b=random_strings(L, n, cb, s0 + 2)))
    return out


def multiroot_requests(count: int, seed: int, scale: float = 0.1, jitter: float = 0.3,
                       workload: str = "sci48open", hf: bool = True):
    """Concurrent multi-root requests on one workload's integrals: request q
    draws its alpha and beta string counts as the workload's counts x scale x
    (1 + jitter u), u uniform in [-1, 1), random weight-n strings from a
    per-request seed, plus (hf) every single excitation of the Hartree-Fock
    string, so that each basis holds a connected HF component."""
    w = WORKLOADS[workload]
    L = w["norb"]
    I0, h, eri = make_integrals(L, w["seed"], w["rank"])
    rng = SplitMix64(seed)
    out = []
    for q in range(count):
        r = rng.draw(4)
        u = (r >> np.uint64(11)).astype(np.float64) / 9007199254740992.0
        ca = max(2, int(round(w["ca"] * scale * (1.0 + jitter * (2.0 * u[0] - 1.0)))))
        cb = max(2, int(round(w["cb"] * scale * (1.0 + jitter * (2.0 * u[1] - 1.0)))))
        s0 = int(r[2] >> np.uint64(1))
        a = random_strings(L, w["na"], ca, s0 + 1)
        b = random_strings(L, w["nb"], cb, s0 + 2)
        if hf:
            a, b = (np.array(sorted(set(x.tolist()) | set(hf_singles(L, n).tolist())),
                             dtype=np.uint64) for x, n in ((a, w["na"]), (b, w["nb"])))
        out.append(dict(id=q, norb=L, I0=I0, h=h, eri=eri, a=a, b=b))
    return out
