"""sbd_amd: MI355X tensor-product-basis (TPB) Davidson and sigma build.

Python host mirror of the reference's TPB operator interface, calling the HIP
library ``libsbd_tpb.so`` through its C ABI (include/sbd_tpb.h).  Names follow
the reference:

* ``SBD``               -- options, sbd::tpb::SBD (chemistry/tpb/sbdiag.h:24-54)
* ``LoadFCIDump``       -- framework/fcidump.h:29-95 (parsed by the library)
* ``LoadAlphaDets``     -- chemistry/basic/makedeterminants.h:167-191
* ``MultTPB``           -- the MultBase<ElemT> operator (chemistry/basic/mult.h:20-103):
                           ``run(hii, Wk, Wb)`` accumulates Wb += H Wk on device
* ``diag``              -- sbd::tpb::diag, method 0 (chemistry/tpb/sbdiag.h:151-526)

Device vectors are torch tensors on the HIP device (torch is used only for
device memory and streams).  There is no CPU fallback: if the library is
missing, loading fails with an error.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsbd_tpb.so")
_LIB = None

EXPORTS = [
    "sbd_status_string", "sbd_last_error", "sbd_options_default", "sbd_pack_strings",
    "sbd_fcidump_read", "sbd_fcidump_entries", "sbd_ctx_create_fcidump", "sbd_ctx_create",
    "sbd_ctx_destroy", "sbd_ctx_dims", "sbd_set_sigma_kernel", "sbd_set_sigma_slab",
    "sbd_make_helpers",
    "sbd_table_counts", "sbd_table_get",
    "sbd_hdiag", "sbd_hdiag_f32", "sbd_sigma", "sbd_sigma_f32", "sbd_inner", "sbd_inner_noalloc",
    "sbd_davidson",
    "sbd_lanczos", "sbd_string_weights", "sbd_occupation", "sbd_correlation", "sbd_carryover",
    "sbd_diag_from", "sbd_diag_from_det", "sbd_save_matrix_form_wf",
    "sbd_save_wavefunction", "sbd_load_wavefunction", "sbd_singles_extend",
    "sbd_singles_extend_strings", "sbd_sort_strings", "sbd_decode_strings_bin",
    "sbd_encode_strings_bin",
    "sbd_diag", "sbd_diag_roots", "sbd_diag_footprint", "sbd_server_create", "sbd_server_submit",
    "sbd_server_wait", "sbd_server_poll", "sbd_server_now", "sbd_server_stats",
    "sbd_server_set_policy",
    "sbd_server_destroy",
]


class SbdError(RuntimeError):
    pass


class Options(C.Structure):
    _fields_ = [("method", C.c_int), ("max_it", C.c_int), ("max_nb", C.c_int),
                ("eps", C.c_double), ("max_time", C.c_double), ("init", C.c_int),
                ("seed", C.c_uint64), ("precision", C.c_int), ("fp32_eps", C.c_double),
                ("verbose", C.c_int)]


class Report(C.Structure):
    _fields_ = [("steps", C.c_int), ("converged", C.c_int), ("residual", C.c_double),
                ("energy", C.c_double), ("t_tables", C.c_double), ("t_hdiag", C.c_double),
                ("t_davidson", C.c_double), ("t_sigma", C.c_double), ("steps_fp32", C.c_int),
                ("nb_used", C.c_int), ("nb_used_fp32", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


MAX_ROOTS = 13


class Request(C.Structure):
    _fields_ = [("norb", C.c_int), ("I0", C.c_double), ("h1", C.POINTER(C.c_double)),
                ("eri_packed", C.POINTER(C.c_double)), ("a_strings", C.POINTER(C.c_uint64)),
                ("na", C.c_size_t), ("b_strings", C.POINTER(C.c_uint64)), ("nb", C.c_size_t),
                ("opt", Options), ("nroots", C.c_int)]


class Result(C.Structure):
    _fields_ = [("energy", C.c_double), ("residual", C.c_double), ("converged", C.c_int),
                ("steps", C.c_int), ("status", C.c_int), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("nroots", C.c_int),
                ("energies", C.c_double * MAX_ROOTS), ("steps_fp32", C.c_int)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["energies"] = list(self.energies)[:max(1, self.nroots)]
        return d


def lib():
    """Load libsbd_tpb.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise SbdError(f"{LIB_PATH} not built: run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp, dp, fp, up, ip, szp = (C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_float),
                               C.POINTER(C.c_uint64), C.POINTER(C.c_int), C.POINTER(C.c_size_t))
    L.sbd_status_string.restype = C.c_char_p
    L.sbd_status_string.argtypes = [C.c_int]
    L.sbd_last_error.restype = C.c_char_p
    L.sbd_last_error.argtypes = [vp]
    L.sbd_options_default.argtypes = [C.POINTER(Options)]
    L.sbd_pack_strings.argtypes = [szp, C.c_size_t, C.c_size_t, C.c_int, up]
    L.sbd_fcidump_read.argtypes = [C.c_char_p, ip, ip, szp]
    L.sbd_fcidump_entries.argtypes = [C.c_char_p, dp, ip, C.c_size_t]
    L.sbd_ctx_create_fcidump.restype = vp
    L.sbd_ctx_create_fcidump.argtypes = [C.c_int, C.c_size_t, dp, ip, up, C.c_size_t, up,
                                         C.c_size_t, ip]
    L.sbd_ctx_create.restype = vp
    L.sbd_ctx_create.argtypes = [C.c_int, C.c_double, dp, dp, up, C.c_size_t, up, C.c_size_t, ip]
    L.sbd_ctx_destroy.argtypes = [vp]
    L.sbd_ctx_dims.argtypes = [vp, up]
    L.sbd_set_sigma_kernel.argtypes = [vp, C.c_int]
    L.sbd_set_sigma_slab.argtypes = [vp, C.c_int]
    L.sbd_make_helpers.argtypes = [vp]
    L.sbd_table_counts.argtypes = [vp, up]
    L.sbd_table_get.argtypes = [vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                C.POINTER(C.c_int32)]
    L.sbd_hdiag.argtypes = [vp, vp, vp]
    L.sbd_hdiag_f32.argtypes = [vp, vp, vp]
    L.sbd_sigma.argtypes = [vp, vp, vp, vp, C.c_int, vp]
    L.sbd_sigma_f32.argtypes = [vp, vp, vp, vp, C.c_int, vp]
    L.sbd_inner.argtypes = [vp, vp, vp, C.c_uint64, dp, vp]
    L.sbd_inner_noalloc.argtypes = [vp, vp, vp, C.c_uint64, dp, vp]
    L.sbd_davidson.argtypes = [vp, vp, vp, C.POINTER(Options), dp, dp, C.POINTER(Report)]
    L.sbd_lanczos.argtypes = [vp, vp, vp, C.POINTER(Options), dp, dp, C.POINTER(Report)]
    L.sbd_string_weights.argtypes = [vp, vp, dp, dp, vp]
    L.sbd_singles_extend.argtypes = [vp, vp, C.c_double, up, C.c_size_t, szp, up, C.c_size_t,
                                     szp, dp, vp]
    L.sbd_singles_extend_strings.argtypes = [C.c_int, up, C.c_size_t, up, C.c_size_t, szp]
    L.sbd_sort_strings.argtypes = [up, C.c_size_t, szp]
    L.sbd_decode_strings_bin.argtypes = [C.c_char_p, C.c_int, up, C.c_size_t, szp]
    L.sbd_encode_strings_bin.argtypes = [C.c_char_p, C.c_int, up, C.c_size_t]
    L.sbd_diag_from.argtypes = [vp, C.POINTER(Options), dp, dp, dp, dp, C.c_int,
                                C.POINTER(Report)]
    L.sbd_diag_from_det.argtypes = [vp, C.POINTER(Options), C.c_uint64, C.c_uint64, dp, dp, dp,
                                    C.c_int, C.POINTER(Report)]
    L.sbd_save_matrix_form_wf.argtypes = [C.c_char_p, C.c_int, up, C.c_size_t, up, C.c_size_t,
                                          dp]
    L.sbd_save_wavefunction.argtypes = [vp, C.c_char_p, C.c_size_t, dp]
    L.sbd_load_wavefunction.argtypes = [vp, C.c_char_p, C.c_size_t, dp, up]
    L.sbd_occupation.argtypes = [vp, vp, dp, vp]
    L.sbd_correlation.argtypes = [vp, vp, dp, dp, vp]
    L.sbd_carryover.argtypes = [vp, vp, C.c_double, C.c_double, up, szp, up, szp, dp, vp]
    L.sbd_diag.argtypes = [vp, C.POINTER(Options), dp, dp, dp, C.c_int, C.POINTER(Report)]
    L.sbd_diag_footprint.argtypes = [vp, C.POINTER(Options), up]
    L.sbd_diag_roots.argtypes = [vp, C.POINTER(Options), C.c_int, dp, dp, dp, dp, C.c_int,
                                 C.POINTER(Report)]
    L.sbd_server_create.restype = vp
    L.sbd_server_create.argtypes = [C.c_uint64, C.c_int, ip]
    L.sbd_server_submit.argtypes = [vp, C.POINTER(Request), up]
    L.sbd_server_wait.argtypes = [vp, C.c_uint64, C.POINTER(Result)]
    L.sbd_server_poll.argtypes = [vp, C.c_uint64, C.POINTER(Result), ip]
    L.sbd_server_now.restype = C.c_double
    L.sbd_server_now.argtypes = [vp]
    L.sbd_server_stats.argtypes = [vp, up]
    L.sbd_server_set_policy.argtypes = [vp, C.c_int, C.c_double]
    L.sbd_server_destroy.argtypes = [vp]
    _LIB = L
    return L


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _dptr(t):
    """Device pointer of a torch tensor (must be contiguous, on the GPU)."""
    if not t.is_cuda or not t.is_contiguous():
        raise SbdError("expected a contiguous device tensor")
    return C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


@dataclass
class SBD:
    """sbd::tpb::SBD (chemistry/tpb/sbdiag.h:24-54), one-rank fields."""
    method: int = 0
    max_it: int = 1
    max_nb: int = 10
    eps: float = 1.0e-4
    max_time: float = 86400.0
    init: int = 0
    seed: int = 1729
    initial_adeterminant_bitstring: str = ""  # explicit initial determinant (alpha string)
    initial_bdeterminant_bitstring: str = ""  # its beta string (default: the alpha one)
    bit_length: int = 20
    dump_matrix_form_wf: str = ""  # SaveMatrixFormWF file (.bin / .dat / .txt)
    do_rdm: int = 0              # 0: occupation density; else also the full 1- and 2-RDMs
    carryover_type: int = 0      # 0 none, 1 dominant strings, 2 + singles, 3 singles of W^2 > threshold
    ratio: float = 0.0
    threshold: float = 0.01
    precision: int = 0           # added: 1 = FP32 vectors + FP64 correction
    fp32_eps: float = 1.0e-2     # added: FP32 -> FP64 hand-over residual
    verbose: int = 0

    def options(self) -> Options:
        o = Options()
        lib().sbd_options_default(C.byref(o))
        for k in ("method", "max_it", "max_nb", "eps", "max_time", "init", "seed", "precision",
                  "fp32_eps", "verbose"):
            setattr(o, k, getattr(self, k))
        return o


def _c_number(text: str, integer: bool):
    """The number C's atoi / atof reads at the start of text (0 if none)."""
    import re
    pat = r"\s*[+-]?\d+" if integer else r"\s*[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?"
    m = re.match(pat, text)
    if not m:
        return 0 if integer else 0.0
    return int(m.group(0)) if integer else float(m.group(0))


def generate_sbd_data(argv) -> SBD:
    """generate_sbd_data (chemistry/tpb/sbdiag.h:56-134): the example programs'
    flags into an SBD, scanned as the reference scans them (the flags tested in
    its order against the current argument, each flag taking the next argument,
    read as C's atoi / atof would; unknown arguments are skipped).  The
    reference's Thrust-only flags and --shuffle are accepted and ignored; the
    communicator sizes must be 1.  A flag without its value raises ValueError."""
    d = SBD()
    comm = {}
    # (flags, kind, field): the reference's order
    table = [(("--adet_comm_size",), "comm", None), (("--bdet_comm_size",), "comm", None),
             (("--task_comm_size",), "comm", None), (("--method",), "i", "method"),
             (("--init",), "i", "init"),
             (("--initial_adeterminant_bitstring", "--initial-adeterminant-bitstring"), "s",
              "initial_adeterminant_bitstring"),
             (("--initial_bdeterminant_bitstring", "--initial-bdeterminant-bitstring"), "s",
              "initial_bdeterminant_bitstring"),
             (("--seed",), "i", "seed"), (("--iteration",), "i", "max_it"),
             (("--block",), "i", "max_nb"), (("--tolerance",), "f", "eps"),
             (("--carryover_type",), "i", "carryover_type"),
             (("--carryover_ratio",), "f", "ratio"), (("--carryover_threshold",), "f", "threshold"),
             (("--max_time",), "f", "max_time"), (("--shuffle",), "skip", None),
             (("--rdm",), "i", "do_rdm"), (("--bit_length",), "i", "bit_length"),
             (("--dump_matrix_form_wf",), "s", "dump_matrix_form_wf"),
             (("--use_precalculated_dets", "--max_memory_gb_for_determinants",
               "--thrust_collapse_loops"), "skip", None)]
    argv = list(argv)
    i = 0
    while i < len(argv):
        for flags, kind, field in table:
            if argv[i] not in flags:
                continue
            if i + 1 >= len(argv):
                raise ValueError(f"{argv[i]} needs a value")
            key = argv[i]
            i += 1
            v = argv[i]
            if kind == "i":
                setattr(d, field, _c_number(v, True))
            elif kind == "f":
                setattr(d, field, _c_number(v, False))
            elif kind == "s":
                setattr(d, field, v)
            elif kind == "comm":
                comm[key] = _c_number(v, True)
        i += 1
    for key, n in comm.items():
        if n != 1:
            raise SbdError(f"{key} {n}: one rank per context (communicator sizes must be 1)")
    return d


def _check(rc, ctx=None):
    if rc != 0:
        msg = lib().sbd_status_string(rc).decode()
        if ctx is not None:
            detail = lib().sbd_last_error(ctx).decode()
            if detail:
                msg += ": " + detail
        raise SbdError(msg)


@dataclass
class FCIDump:
    """sbd::FCIDump: header key/values and (value, i, j, k, l) entries in file order."""
    header: dict = field(default_factory=dict)
    values: np.ndarray = None
    idx: np.ndarray = None

    @property
    def norb(self):
        return int(self.header["NORB"])

    @property
    def nelec(self):
        return int(self.header.get("NELEC", 0))


def LoadFCIDump(path: str) -> FCIDump:
    norb, nelec, n = C.c_int(), C.c_int(), C.c_size_t()
    _check(lib().sbd_fcidump_read(path.encode(), C.byref(norb), C.byref(nelec), C.byref(n)))
    vals = np.zeros(n.value)
    idx = np.zeros(4 * n.value, dtype=np.int32)
    _check(lib().sbd_fcidump_entries(path.encode(), _ptr(vals, C.c_double), _ptr(idx, C.c_int),
                                     n.value))
    return FCIDump({"NORB": str(norb.value), "NELEC": str(nelec.value)}, vals, idx.reshape(-1, 4))


def from_string(s: str, L: int) -> int:
    """framework/bit_manipulation.h:219-235: character s[L-1-i] is orbital i."""
    v = 0
    for i in range(L):
        if s[L - 1 - i] == "1":
            v |= 1 << i
    return v


def from_string_checked(bitstring: str, bit_length: int, total_bits: int) -> int:
    """framework/determinant_initialization.h:18-34: a bit string of exactly
    total_bits characters '0'/'1' (ValueError otherwise), as one integer
    (orbital p -> bit p)."""
    if bit_length == 0 or bit_length > 64:
        raise ValueError("invalid determinant bit length")
    if len(bitstring) != total_bits:
        raise ValueError("determinant bitstring length does not match the basis")
    if any(ch not in "01" for ch in bitstring):
        raise ValueError("determinant bitstring must contain only 0 and 1")
    return from_string(bitstring, total_bits)


def save_matrix_form_wf(filename: str, norb: int, a, b, W):
    """SaveMatrixFormWF (tpb/restart.h:663-780) at one rank: by extension,
    "bin" writes W's doubles, "dat"/"txt" one line per determinant
    ("<W> # <ia>: <alpha bits> <ib>: <beta bits>"), others nothing."""
    a, b = _u64(a), _u64(b)
    W = np.ascontiguousarray(W, dtype=np.float64)
    if W.size != len(a) * len(b):
        raise SbdError("W must have len(a) * len(b) elements")
    _check(lib().sbd_save_matrix_form_wf(str(filename).encode(), int(norb), _ptr(a, C.c_uint64),
                                         len(a), _ptr(b, C.c_uint64), len(b),
                                         _ptr(W, C.c_double)))


def LoadAlphaDets(path: str, L: int) -> np.ndarray:
    """Text string list -> sorted unique uint64 strings (makedeterminants.h:167-191)."""
    with open(path) as f:
        vals = {from_string(line.rstrip("\n"), L) for line in f if line.strip()}
    return np.array(sorted(vals), dtype=np.uint64)


def pack_strings(words: np.ndarray, bit_length: int, norb: int) -> np.ndarray:
    """Reference det_vector rows (size_t words of bit_length bits) -> uint64 strings."""
    words = np.ascontiguousarray(words, dtype=np.uint64)
    n = words.shape[0]
    out = np.zeros(n, dtype=np.uint64)
    _check(lib().sbd_pack_strings(words.ctypes.data_as(C.POINTER(C.c_size_t)), n, bit_length, norb,
                                  _ptr(out, C.c_uint64)))
    return out


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def sort_bitarray(strings):
    """sort_bitarray (bit_manipulation.h:421-449): ascending, unique."""
    s = _u64(strings).copy()
    n = C.c_size_t(0)
    _check(lib().sbd_sort_strings(_ptr(s, C.c_uint64), len(s), C.byref(n)))
    return s[:n.value]


def singles_extend_strings(strings, norb):
    """SinglesExtendHalfdets, string form (extend.h:186-324): the strings and
    all their single excitations, sorted and unique."""
    s = _u64(strings)
    cap = max(1, len(s) * (1 + norb * norb // 4))
    out = np.zeros(cap, dtype=np.uint64)
    n = C.c_size_t(0)
    _check(lib().sbd_singles_extend_strings(int(norb), _ptr(s, C.c_uint64), len(s),
                                            _ptr(out, C.c_uint64), cap, C.byref(n)))
    return out[:n.value].copy()


def DecodeAlphaDets(path, norb):
    """DecodeAlphaDets (makedeterminants.h:61-114): strings of a binary file in
    file order (unsorted; the CLI sorts them with sort_bitarray)."""
    n = C.c_size_t(0)
    rc = lib().sbd_decode_strings_bin(str(path).encode(), int(norb), None, 0, C.byref(n))
    if rc not in (0, 1):
        _check(rc)
    out = np.zeros(max(1, n.value), dtype=np.uint64)
    _check(lib().sbd_decode_strings_bin(str(path).encode(), int(norb), _ptr(out, C.c_uint64),
                                        len(out), C.byref(n)))
    return out[:n.value].copy()


def encode_strings_bin(path, strings, norb):
    """The carryover.bin writer (main.cc:245-260)."""
    s = _u64(strings)
    _check(lib().sbd_encode_strings_bin(str(path).encode(), int(norb), _ptr(s, C.c_uint64),
                                        len(s)))


class MultTPB:
    """TPB Hamiltonian operator on one GPU (MultBase / MultTPBThrust equivalent).

    Owns the device copies of strings, integrals and excitation tables."""

    def __init__(self, norb, a, b, *, fcidump: FCIDump | None = None, I0=0.0, h=None, eri=None,
                 build_tables=True):
        self.norb = int(norb)
        self.a = np.ascontiguousarray(a, dtype=np.uint64)
        self.b = np.ascontiguousarray(b, dtype=np.uint64)
        st = C.c_int(0)
        if fcidump is not None:
            vals = np.ascontiguousarray(fcidump.values, dtype=np.float64)
            idx = np.ascontiguousarray(fcidump.idx, dtype=np.int32).reshape(-1)
            h_ = lib().sbd_ctx_create_fcidump(self.norb, len(vals), _ptr(vals, C.c_double),
                                              _ptr(idx, C.c_int), _ptr(self.a, C.c_uint64),
                                              len(self.a), _ptr(self.b, C.c_uint64), len(self.b),
                                              C.byref(st))
        else:
            h = np.ascontiguousarray(h, dtype=np.float64)
            eri = np.ascontiguousarray(eri, dtype=np.float64)
            h_ = lib().sbd_ctx_create(self.norb, float(I0), _ptr(h, C.c_double),
                                      _ptr(eri, C.c_double), _ptr(self.a, C.c_uint64), len(self.a),
                                      _ptr(self.b, C.c_uint64), len(self.b), C.byref(st))
        if not h_:
            _check(st.value if st.value else 1)
        self._ctx = C.c_void_p(h_)
        self.na, self.nb = len(self.a), len(self.b)
        self.N = self.na * self.nb
        if build_tables:
            self.make_helpers()

    def close(self):
        if getattr(self, "_ctx", None):
            lib().sbd_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: module globals may be gone
            pass

    def make_helpers(self):
        _check(lib().sbd_make_helpers(self._ctx), self._ctx)

    def set_sigma_kernel(self, kernel: str = "auto"):
        """Sigma kernel: "auto" (by sparsity), "rows" (row kernel; dense
        alpha-single x beta-single terms in a pass of their own), "elements"
        (one element per thread), "rows_inline" (row kernel, every term
        inline) or "split" (a streaming kernel for the diagonal and alpha
        terms over column slabs, then a row kernel for the beta terms); all
        give the same results."""
        k = {"auto": 0, "rows": 1, "elements": 2, "rows_inline": 3, "split": 4}[kernel]
        _check(lib().sbd_set_sigma_kernel(self._ctx, k), self._ctx)

    def set_sigma_slab(self, cols: int):
        """Slab width (256, 512 or 1024 columns) of the split sigma's streaming
        kernel; results do not depend on it."""
        _check(lib().sbd_set_sigma_slab(self._ctx, int(cols)), self._ctx)

    def table_counts(self):
        out = np.zeros(4, dtype=np.uint64)
        _check(lib().sbd_table_counts(self._ctx, _ptr(out, C.c_uint64)), self._ctx)
        return out.astype(np.int64)

    def table(self, which):
        """(offsets, kets, CrAn) of table `which` in the reference layout."""
        n = self.na if which < 2 else self.nb
        cnt = int(self.table_counts()[which])
        w = 2 if which % 2 == 0 else 4
        off = np.zeros(n + 1, dtype=np.int64)
        ket = np.zeros(max(cnt, 1), dtype=np.int64)
        cr = np.zeros(max(w * cnt, 1), dtype=np.int32)
        _check(lib().sbd_table_get(self._ctx, which, _ptr(off, C.c_int64), _ptr(ket, C.c_int64),
                                   _ptr(cr, C.c_int32)), self._ctx)
        return off, ket[:cnt], cr[:w * cnt].reshape(cnt, w)

    def hdiag(self, dtype=None):
        import torch
        dtype = dtype or torch.float64
        out = torch.empty(self.N, dtype=dtype, device="cuda")
        fn = lib().sbd_hdiag if dtype == torch.float64 else lib().sbd_hdiag_f32
        _check(fn(self._ctx, _dptr(out), _stream()), self._ctx)
        return out

    # alias matching MultTPBThrust::makeQChamDiagTerms
    makeQChamDiagTerms = hdiag

    def run(self, hii, Wk, Wb, accumulate=True):
        """MultBase::run: Wb += H Wk (device tensors, float64 or float32)."""
        import torch
        if not (hii.dtype == Wk.dtype == Wb.dtype) or Wk.numel() != self.N or Wb.numel() != self.N \
                or hii.numel() != self.N:
            raise SbdError("run: shapes/dtypes must match the basis")
        fn = lib().sbd_sigma if Wk.dtype == torch.float64 else lib().sbd_sigma_f32
        _check(fn(self._ctx, _dptr(hii), _dptr(Wk), _dptr(Wb), 1 if accumulate else 0, _stream()),
               self._ctx)
        return Wb

    def sigma(self, hii, Wk):
        import torch
        Wb = torch.empty_like(Wk)
        return self.run(hii, Wk, Wb, accumulate=False)

    def inner(self, x, y):
        out = C.c_double(0.0)
        _check(lib().sbd_inner(self._ctx, _dptr(x), _dptr(y), x.numel(), C.byref(out), _stream()),
               self._ctx)
        return out.value

    def davidson(self, hii, W, sbd: SBD, history=False):
        """Davidson on device vectors; W (device, float64) is updated in place."""
        o = sbd.options()
        E = C.c_double(0.0)
        rep = Report()
        hist = np.zeros(4 * sbd.max_it * sbd.max_nb)
        _check(lib().sbd_davidson(self._ctx, _dptr(hii), _dptr(W), C.byref(o), C.byref(E),
                                  _ptr(hist, C.c_double), C.byref(rep)), self._ctx)
        out = (E.value, rep.as_dict())
        if history:
            out = out + (hist[:4 * rep.steps].reshape(-1, 4),)
        return out

    def lanczos(self, hii, W, sbd: SBD, history=False):
        """Restarted Lanczos (sbd::Lanczos, lanczos.h:162-327) on device vectors;
        W (device, float64) is updated in place to sum U[ib, 0] C[ib].  History
        rows are (E0, beta, seconds, 64)."""
        o = sbd.options()
        E = C.c_double(0.0)
        rep = Report()
        hist = np.zeros(4 * sbd.max_it * sbd.max_nb)
        _check(lib().sbd_lanczos(self._ctx, _dptr(hii), _dptr(W), C.byref(o), C.byref(E),
                                 _ptr(hist, C.c_double), C.byref(rep)), self._ctx)
        out = (E.value, rep.as_dict())
        if history:
            out = out + (hist[:4 * rep.steps].reshape(-1, 4),)
        return out

    def save_wavefunction(self, prefix, W, bit_length=20):
        """SaveWavefunction (restart.h:20-65): writes prefix + '000000'."""
        W = np.ascontiguousarray(W, dtype=np.float64)
        _check(lib().sbd_save_wavefunction(self._ctx, str(prefix).encode(), int(bit_length),
                                           _ptr(W, C.c_double)), self._ctx)

    def load_wavefunction(self, prefix, bit_length=20):
        """LoadWavefunction (restart.h:414-657): the saved coefficients mapped
        onto this basis and normalised; returns (W, number placed)."""
        W = np.zeros(self.N)
        placed = C.c_uint64(0)
        _check(lib().sbd_load_wavefunction(self._ctx, str(prefix).encode(), int(bit_length),
                                           _ptr(W, C.c_double), C.byref(placed)), self._ctx)
        return W, placed.value

    def diag(self, sbd: SBD, want_vector=False, history=False, W_init=None, start_det=None):
        """sbd::tpb::diag: method 0/1 Davidson, 2/3 Lanczos.  Returns (energy,
        W or None, report[, history]); Davidson history rows are (E0, |R|,
        seconds since solver start, vector precision 32/64), Lanczos rows
        (E0, beta, seconds, 64).  W_init: start vector (host, N); start_det:
        (alpha string, beta string) whose unit vector starts the solve
        (BasisInitVectorFromDeterminants)."""
        o = sbd.options()
        E = C.c_double(0.0)
        rep = Report()
        W = np.zeros(self.N) if want_vector else None
        cap = sbd.max_it * sbd.max_nb * (2 if sbd.precision == 1 else 1)
        hist = np.zeros(4 * cap)
        if start_det is not None:
            _check(lib().sbd_diag_from_det(self._ctx, C.byref(o), C.c_uint64(int(start_det[0])),
                                           C.c_uint64(int(start_det[1])), C.byref(E),
                                           _ptr(W, C.c_double) if want_vector else None,
                                           _ptr(hist, C.c_double), cap, C.byref(rep)), self._ctx)
        elif W_init is not None:
            W0 = np.ascontiguousarray(W_init, dtype=np.float64)
            if W0.size != self.N:
                raise SbdError("W_init must have N elements")
            _check(lib().sbd_diag_from(self._ctx, C.byref(o), _ptr(W0, C.c_double), C.byref(E),
                                       _ptr(W, C.c_double) if want_vector else None,
                                       _ptr(hist, C.c_double), cap, C.byref(rep)), self._ctx)
        else:
            _check(lib().sbd_diag(self._ctx, C.byref(o), C.byref(E),
                                  _ptr(W, C.c_double) if want_vector else None,
                                  _ptr(hist, C.c_double), cap, C.byref(rep)), self._ctx)
        out = (E.value, W, rep.as_dict())
        if history:
            out = out + (hist[:4 * min(rep.steps, cap)].reshape(-1, 4),)
        return out

    def diag_roots(self, sbd: SBD, nroots: int, want_vectors=False, history=False):
        """Block Davidson for the nroots lowest states: returns (energies,
        residuals, report[, W (nroots, N)][, history rows (E0, max |R|, s, 32/64)])."""
        o = sbd.options()
        E = np.zeros(nroots)
        res = np.zeros(nroots)
        W = np.zeros((nroots, self.N)) if want_vectors else None
        cap = (o.max_it + 1) * 2 if history else 0
        H = np.zeros(4 * max(cap, 1))
        rep = Report()
        _check(lib().sbd_diag_roots(self._ctx, C.byref(o), int(nroots), _ptr(E, C.c_double),
                                    _ptr(res, C.c_double),
                                    _ptr(W, C.c_double) if W is not None else None,
                                    _ptr(H, C.c_double) if history else None, cap,
                                    C.byref(rep)), self._ctx)
        out = [E, res, rep.as_dict()]
        if want_vectors:
            out.append(W)
        if history:
            h = H.reshape(-1, 4)
            out.append(h[: int(np.count_nonzero(h[:, 3]))])
        return tuple(out)

    def string_weights(self, W):
        """(alpha, beta) string weights of device vector W (rdmat.h:11-58, 136-182)."""
        wa, wb = np.zeros(self.na), np.zeros(self.nb)
        _check(lib().sbd_string_weights(self._ctx, _dptr(W), _ptr(wa, C.c_double),
                                        _ptr(wb, C.c_double), _stream()), self._ctx)
        return wa, wb

    def occupation(self, W):
        """OccupationDensity (occupation.h:10-95): density[2o] alpha, [2o+1] beta."""
        out = np.zeros(2 * self.norb)
        _check(lib().sbd_occupation(self._ctx, _dptr(W), _ptr(out, C.c_double), _stream()),
               self._ctx)
        return out

    def correlation(self, W):
        """Correlation (correlation.h:379-643, do_rdm != 0): (onebody, twobody) as
        the reference's one_p_rdm / two_p_rdm, arrays of shape (2, L*L) and
        (4, L**4): onebody[s][o1 + L*o2], twobody[s1 + 2*s2][o1 + L*o2 + L^2*o3 + L^3*o4]."""
        L = self.norb
        one = np.zeros((2, L * L))
        two = np.zeros((4, L ** 4))
        _check(lib().sbd_correlation(self._ctx, _dptr(W), _ptr(one, C.c_double),
                                     _ptr(two, C.c_double), _stream()), self._ctx)
        return one, two

    def carryover(self, W, threshold=0.01, ratio=0.0):
        """CarryOverAdet/Bdet (rdmat.h:60-134): (alpha strings, beta strings,
        truncated weights) in descending order of weight."""
        ao, bo = np.zeros(self.na, dtype=np.uint64), np.zeros(self.nb, dtype=np.uint64)
        na, nb = C.c_size_t(0), C.c_size_t(0)
        tr = np.zeros(2)
        _check(lib().sbd_carryover(self._ctx, _dptr(W), float(threshold), float(ratio),
                                   _ptr(ao, C.c_uint64), C.byref(na), _ptr(bo, C.c_uint64),
                                   C.byref(nb), _ptr(tr, C.c_double), _stream()), self._ctx)
        return ao[:na.value].copy(), bo[:nb.value].copy(), tr

    def singles_extend(self, W, cutoff):
        """SinglesExtendHalfdets, amplitude form (extend.h:14-185): (alpha
        strings, beta strings, weight of the determinants above the cutoff)."""
        L = self.norb
        ca = max(1, self.na * (1 + L * L // 4))
        cb = max(1, self.nb * (1 + L * L // 4))
        ao, bo = np.zeros(ca, dtype=np.uint64), np.zeros(cb, dtype=np.uint64)
        na, nb = C.c_size_t(0), C.c_size_t(0)
        tw = C.c_double(0.0)
        _check(lib().sbd_singles_extend(self._ctx, _dptr(W), float(cutoff), _ptr(ao, C.c_uint64),
                                        ca, C.byref(na), _ptr(bo, C.c_uint64), cb, C.byref(nb),
                                        C.byref(tw), _stream()), self._ctx)
        return ao[:na.value].copy(), bo[:nb.value].copy(), tw.value

    def footprint(self, sbd: SBD):
        o = sbd.options()
        out = C.c_uint64(0)
        _check(lib().sbd_diag_footprint(self._ctx, C.byref(o), C.byref(out)), self._ctx)
        return out.value


@dataclass
class DiagResult:
    """Outputs of sbd::tpb::diag (chemistry/tpb/sbdiag.h:151-163): energy,
    occupation density (density[2o] alpha, [2o+1] beta), carry-over strings
    (empty unless carryover_type is 1..3), the final vector and, for do_rdm != 0,
    the one- and two-particle RDMs (empty otherwise)."""
    energy: float
    density: np.ndarray
    co_adet: np.ndarray
    co_bdet: np.ndarray
    W: np.ndarray
    one_p_rdm: np.ndarray = None
    two_p_rdm: np.ndarray = None


def diag(sbd_data: SBD, fcidump: FCIDump, adet, bdet, loadname="", savename=""):
    """sbd::tpb::diag (one rank; method 0/1 Davidson, 2/3 Lanczos), followed by
    the occupation density (do_rdm == 0) or the full one- and two-particle RDMs
    with the density read off the one-particle diagonal (do_rdm != 0), and the
    carry-over strings (sbdiag.h:666-791).  Returns a DiagResult."""
    import torch
    if sbd_data.carryover_type not in (0, 1, 2, 3):
        raise SbdError("carryover_type must be 0..3")
    start = None
    if not loadname:
        # explicit initial determinant (sbdiag.h:303-335), checked as
        # BasisInitVectorFromDeterminants checks it (davidson.h:97-110)
        sa, sb = sbd_data.initial_adeterminant_bitstring, sbd_data.initial_bdeterminant_bitstring
        if sb and not sa:
            raise ValueError("an initial beta determinant requires an initial alpha determinant")
        if sa:
            L = fcidump.norb
            want = (from_string_checked(sa, sbd_data.bit_length, L),
                    from_string_checked(sb or sa, sbd_data.bit_length, L))
            for name, basis, x in (("alpha", adet, want[0]), ("beta", bdet, want[1])):
                hits = int(np.count_nonzero(np.asarray(basis, dtype=np.uint64) == np.uint64(x)))
                if hits == 0:
                    raise ValueError(f"initial {name} determinant is not present in the basis")
                if hits > 1:
                    raise ValueError(f"initial {name} determinant occurs more than once in the "
                                     "basis")
            start = want
    m = MultTPB(fcidump.norb, adet, bdet, fcidump=fcidump)
    try:
        W0 = None
        if loadname:
            # initial vector from a saved wavefunction (sbdiag.h:339-347)
            W0, placed = m.load_wavefunction(loadname, sbd_data.bit_length)
            if placed == 0:
                raise SbdError(f"{loadname}: no saved coefficient lies in this basis")
        E, W, _ = m.diag(sbd_data, want_vector=True, W_init=W0, start_det=start)
        if savename:
            m.save_wavefunction(savename, W, sbd_data.bit_length)
        if sbd_data.dump_matrix_form_wf:
            # SaveMatrixFormWF (sbdiag.h:808-813)
            save_matrix_form_wf(sbd_data.dump_matrix_form_wf, m.norb, adet, bdet, W)
        Wd = torch.tensor(W, device="cuda")
        one = two = np.zeros((0, 0))
        if sbd_data.do_rdm == 0:
            density = m.occupation(Wd)
        else:
            # sbdiag.h:704-728: density[2o + s] = one_p_rdm[s][o + L*o]
            one, two = m.correlation(Wd)
            L = m.norb
            density = np.zeros(2 * L)
            for o in range(L):
                density[2 * o] = one[0][o + L * o]
                density[2 * o + 1] = one[1][o + L * o]
        co_a = np.zeros(0, dtype=np.uint64)
        co_b = np.zeros(0, dtype=np.uint64)
        # carry-over strings (sbdiag.h:745-791): 1 dominant strings, 2 those
        # extended by singles, 3 singles extension of the strings of the
        # determinants above the threshold
        if sbd_data.carryover_type in (1, 2):
            co_a, co_b, _ = m.carryover(Wd, sbd_data.threshold, sbd_data.ratio)
            if sbd_data.carryover_type == 2:
                co_a = singles_extend_strings(co_a, m.norb)
                co_b = singles_extend_strings(co_b, m.norb)
        elif sbd_data.carryover_type == 3:
            co_a, co_b, _ = m.singles_extend(Wd, sbd_data.threshold)
        del Wd
    finally:
        m.close()
    return DiagResult(E, density, co_a, co_b, W, one, two)


class Server:
    """Continuous-batching solver service on one GPU (sbd_server_* in sbd_tpb.h).

    ``submit`` copies a request's integrals and strings and returns its id at
    once; ``wait`` blocks until that solve is done and returns its result
    (energy, per-root energies, residual, converged, steps, status, and the
    submit / admit / done times on the server clock).  A single-root FP64
    solve's result is bit-identical alone, in any mix, and to ``MultTPB.diag``
    with the same options; FP32 and multi-root requests run whole beside the
    batches and equal ``MultTPB.diag`` / ``MultTPB.diag_roots``.  ``policy`` is the
    admission order: "fifo" (arrival order) or "sjf" (smallest footprint
    first; a request waiting ``age_limit`` seconds goes first)."""

    def __init__(self, mem_budget_bytes: int, max_active: int = 32, policy: str = "fifo",
                 age_limit: float = 30.0):
        if policy not in ("fifo", "sjf"):
            raise SbdError("policy must be 'fifo' or 'sjf'")
        st = C.c_int(0)
        h = lib().sbd_server_create(int(mem_budget_bytes), int(max_active), C.byref(st))
        if not h:
            _check(st.value if st.value else 1)
        self._h = C.c_void_p(h)
        _check(lib().sbd_server_set_policy(self._h, 0 if policy == "fifo" else 1,
                                           float(age_limit)))

    def submit(self, norb, a, b, *, I0=0.0, h=None, eri=None, sbd: "SBD | None" = None,
               nroots: int = 1) -> int:
        """Queue a solve: the lowest root (nroots 1) or the nroots lowest
        (block Davidson); sbd.precision 1 runs the FP32 Davidson + FP64 sweep."""
        sbd = sbd or SBD()
        h = np.ascontiguousarray(h, dtype=np.float64)
        eri = np.ascontiguousarray(eri, dtype=np.float64)
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.uint64)
        rq = Request(int(norb), float(I0), _ptr(h, C.c_double), _ptr(eri, C.c_double),
                     _ptr(a, C.c_uint64), len(a), _ptr(b, C.c_uint64), len(b), sbd.options(),
                     int(nroots))
        rid = C.c_uint64(0)
        _check(lib().sbd_server_submit(self._h, C.byref(rq), C.byref(rid)))
        return rid.value

    def wait(self, rid: int) -> dict:
        r = Result()
        _check(lib().sbd_server_wait(self._h, C.c_uint64(rid), C.byref(r)))
        d = r.as_dict()
        if d["status"] != 0:
            raise SbdError(f"solve {rid} failed: " + lib().sbd_status_string(d["status"]).decode())
        return d

    def poll(self, rid: int):
        r, done = Result(), C.c_int(0)
        _check(lib().sbd_server_poll(self._h, C.c_uint64(rid), C.byref(r), C.byref(done)))
        return r.as_dict() if done.value else None

    def now(self) -> float:
        return lib().sbd_server_now(self._h)

    def stats(self) -> dict:
        out = np.zeros(2, dtype=np.uint64)
        _check(lib().sbd_server_stats(self._h, _ptr(out, C.c_uint64)))
        return {"rounds": int(out[0]), "peak_active": int(out[1])}

    def close(self):
        if getattr(self, "_h", None):
            lib().sbd_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: module globals may be gone
            pass
