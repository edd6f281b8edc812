"""CPU tests of the C ABI library: it loads, exports every function that
include/sbd_tpb.h declares, and its host-only entry points (string packing,
FCIDUMP parsing, argument validation) behave like the reference's host code.
No compute call is made here (there is no GPU in this container)."""
import gzip
import os
import re
import shutil

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


@pytest.fixture(scope="module")
def L():
    import sbd_amd
    return sbd_amd.lib()


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "sbd_tpb.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sbd_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol(L):
    import sbd_amd
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), f"{n} not exported"
    assert sorted(sbd_amd.EXPORTS) == names


def test_pack_strings_kat(L):
    """from_string_checked("10000001", bit_length 4, 8) -> words {1, 8}
    (test_initial_determinant.cc:16-24) packs back to 0b10000001."""
    import sbd_amd
    words = np.array([[1, 8]], dtype=np.uint64)
    assert sbd_amd.pack_strings(words, 4, 8).tolist() == [0b10000001]
    # default bit_length 20, norb 36 (two words per string)
    rng = np.random.default_rng(1)
    vals = [int(x) & ((1 << 36) - 1) for x in rng.integers(0, 2 ** 62, 50, dtype=np.int64)]
    w = np.array([[v & ((1 << 20) - 1), v >> 20] for v in vals], dtype=np.uint64)
    assert sbd_amd.pack_strings(w, 20, 36).tolist() == vals


def test_from_string_matches_reference_rule():
    import sbd_amd
    assert sbd_amd.from_string("000000000000000000011111", 24) == 0b11111
    assert sbd_amd.from_string("100", 3) == 0b100


def test_fcidump_parse_matches_oracle(L, tmp_path):
    import sbd_amd
    from oracle import oracle as O
    src = os.path.join(GOLDEN, "n2", "fcidump.txt.gz")
    plain = tmp_path / "FCIDUMP"
    with gzip.open(src, "rb") as fi, open(plain, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    fd = sbd_amd.LoadFCIDump(str(plain))
    hdr, vals, idx = O.load_fcidump(src)
    assert fd.norb == int(hdr["NORB"]) == 16 or fd.norb == int(hdr["NORB"])
    assert fd.nelec == int(hdr["NELEC"])
    assert np.array_equal(fd.values, vals)
    assert np.array_equal(fd.idx, idx)


def test_missing_file_is_io_error(L):
    import sbd_amd
    with pytest.raises(sbd_amd.SbdError, match="i/o"):
        sbd_amd.LoadFCIDump("/nonexistent/FCIDUMP")


def test_argument_validation_before_device(L):
    """Bad inputs are rejected with SBD_ERR_ARG before any device call."""
    import ctypes as C
    import sbd_amd
    st = C.c_int(-1)
    a = np.array([0b0011, 0b0101], dtype=np.uint64)
    h = np.zeros(16)
    eri = np.zeros(55)
    P = lambda x, t: x.ctypes.data_as(C.POINTER(t))  # noqa: E731
    # norb out of range
    r = L.sbd_ctx_create(65, 0.0, P(h, C.c_double), P(eri, C.c_double), P(a, C.c_uint64), 2,
                         P(a, C.c_uint64), 2, C.byref(st))
    assert not r and st.value == 1
    # unsorted strings (reference lists are sorted, unique)
    bad = np.array([0b0101, 0b0011], dtype=np.uint64)
    r = L.sbd_ctx_create(4, 0.0, P(h, C.c_double), P(eri, C.c_double), P(bad, C.c_uint64), 2,
                         P(a, C.c_uint64), 2, C.byref(st))
    assert not r and st.value == 1
    # string with bits beyond norb
    big = np.array([0b0011, 0b10001], dtype=np.uint64)
    r = L.sbd_ctx_create(4, 0.0, P(h, C.c_double), P(eri, C.c_double), P(big, C.c_uint64), 2,
                         P(a, C.c_uint64), 2, C.byref(st))
    assert not r and st.value == 1
    # FCIDUMP entry with an orbital index > norb
    vals = np.array([1.0])
    idx = np.array([5, 1, 0, 0], dtype=np.int32)
    r = L.sbd_ctx_create_fcidump(4, 1, P(vals, C.c_double), P(idx, C.c_int), P(a, C.c_uint64), 2,
                                 P(a, C.c_uint64), 2, C.byref(st))
    assert not r and st.value == 1
    # status strings
    assert L.sbd_status_string(0) == b"ok"
    assert L.sbd_status_string(4) == b"out of device memory"


def test_options_default_match_reference(L):
    """sbd::tpb::SBD defaults (chemistry/tpb/sbdiag.h:30-45)."""
    import sbd_amd
    o = sbd_amd.SBD().options()
    assert (o.method, o.max_it, o.max_nb, o.init, o.seed) == (0, 1, 10, 0, 1729)
    assert o.eps == 1e-4 and o.max_time == 86400.0


def test_no_cpu_fallback_without_library(monkeypatch, tmp_path):
    """The product path fails loudly when the HIP library is missing."""
    import sbd_amd
    monkeypatch.setattr(sbd_amd, "_LIB", None)
    monkeypatch.setattr(sbd_amd, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(sbd_amd.SbdError, match="not built"):
        sbd_amd.lib()


def _all_strings(L, n):
    from itertools import combinations
    return np.array(sorted(sum(1 << p for p in c) for c in combinations(range(L), n)),
                    dtype=np.uint64)


def test_singles_extension_closed_forms(L):
    """SinglesExtendHalfdets, string form (extend.h:186-324): one string of n
    electrons in L orbitals extends to 1 + n(L-n) strings, exactly those of
    Hamming distance 0 or 2 with the same electron count (brute force); the
    full set of weight-n strings is closed under singles; the output is
    sorted and unique (sort_bitarray)."""
    import sbd_amd
    for Lo, n in ((8, 3), (12, 4), (40, 2)):
        hf = (1 << n) - 1
        out = sbd_amd.singles_extend_strings([hf], Lo)
        assert len(out) == 1 + n * (Lo - n)
        assert np.all(np.diff(out.astype(np.float64)) > 0)
        if Lo <= 12:
            full = _all_strings(Lo, n)
            dist = np.array([bin(int(s) ^ hf).count("1") for s in full])
            assert np.array_equal(out, full[dist <= 2])
            assert np.array_equal(sbd_amd.singles_extend_strings(full, Lo), full)
    assert np.array_equal(sbd_amd.sort_bitarray([5, 3, 5, 9, 3]), np.array([3, 5, 9],
                                                                           dtype=np.uint64))


def test_binary_strings_kats(L, tmp_path):
    """DecodeAlphaDets (makedeterminants.h:61-114): a record of ceil(L/8) bytes
    is a big-endian number whose low L bits are the string (leading padding
    skipped); the carryover.bin writer (main.cc:245-260) shifts the string to
    the top of the record (padding last).  For L = 12, string {0, 5, 11} =
    0x821: decoded from bytes 08 21; written as bytes 82 10."""
    import sbd_amd
    p = tmp_path / "s.bin"
    p.write_bytes(bytes([0x08, 0x21, 0xF0, 0x01]))    # second record: padding bits set
    got = sbd_amd.DecodeAlphaDets(p, 12)
    assert list(got) == [0x821, 0x001]
    sbd_amd.encode_strings_bin(p, [0x821], 12)
    assert p.read_bytes() == bytes([0x82, 0x10])
    rng = np.random.default_rng(4)
    for Lo in (8, 16, 36, 64):
        s = rng.integers(0, 2 ** min(Lo, 63), size=20, dtype=np.uint64)
        nb = (Lo + 7) // 8
        p.write_bytes(b"".join(int(v).to_bytes(nb, "big") for v in s))
        assert np.array_equal(sbd_amd.DecodeAlphaDets(p, Lo), s)
        if Lo % 8 == 0:   # writer and reader agree when there is no padding
            sbd_amd.encode_strings_bin(p, s, Lo)
            assert np.array_equal(sbd_amd.DecodeAlphaDets(p, Lo), s)


def test_from_string_checked_kats():
    """determinant_initialization.h:18-34 with the reference's known answers
    (test_initial_determinant.cc:16-27)."""
    import sbd_amd
    assert sbd_amd.from_string_checked("10000001", 4, 8) == 0b10000001
    for s, bl, n in (("1000000", 4, 8), ("1000000x", 4, 8), ("00101", 0, 5)):
        with pytest.raises(ValueError):
            sbd_amd.from_string_checked(s, bl, n)


def test_save_matrix_form_wf_layouts(L, tmp_path):
    """SaveMatrixFormWF (restart.h:663-780) at one rank: "bin" is W's doubles
    as they lie; "dat"/"txt" (any case; get_extension lower-cases and drops
    spaces) is one line per determinant, value as a default ostream prints a
    double (six significant digits), then "# ia: <alpha bits> ib: <beta bits>"
    with orbital norb - 1 first (makestring); other extensions write nothing."""
    import sbd_amd
    a = np.array([0b011, 0b101], dtype=np.uint64)
    b = np.array([0b011, 0b110], dtype=np.uint64)
    W = np.array([0.5, -1.25e-7, 1.0 / 3.0, 0.0])
    sbd_amd.save_matrix_form_wf(tmp_path / "w.bin", 3, a, b, W)
    assert (tmp_path / "w.bin").read_bytes() == W.astype("<f8").tobytes()
    for name in ("w.txt", "w.DAT", "w.tx t"):
        sbd_amd.save_matrix_form_wf(tmp_path / name, 3, a, b, W)
        assert (tmp_path / name).read_text().splitlines() == [
            "0.5 # 0: 011 0: 011", "-1.25e-07 # 0: 011 1: 110",
            "0.333333 # 1: 101 0: 011", "0 # 1: 101 1: 110"]
    sbd_amd.save_matrix_form_wf(tmp_path / "w.wf", 3, a, b, W)
    sbd_amd.save_matrix_form_wf(tmp_path / "noext", 3, a, b, W)
    assert not (tmp_path / "w.wf").exists() and not (tmp_path / "noext").exists()


def test_generate_sbd_data_flags():
    """generate_sbd_data (sbdiag.h:56-134): every flag of the reference into its
    field, atoi / atof reading, both spellings of the initial-determinant
    flags, Thrust-only flags ignored, defaults kept otherwise."""
    import sbd_amd
    argv = ["prog", "--method", "2", "--init", "1", "--seed", "12x", "--iteration", "20",
            "--block", "8", "--tolerance", "1e-8", "--carryover_type", "3",
            "--carryover_ratio", "0.25", "--carryover_threshold", ".5e-2", "--max_time", "60",
            "--shuffle", "1", "--rdm", "1", "--bit_length", "64",
            "--initial-adeterminant-bitstring", "0011", "--initial_bdeterminant_bitstring",
            "0101", "--dump_matrix_form_wf", "w.txt", "--use_precalculated_dets", "1",
            "--adet_comm_size", "1", "unknown"]
    d = sbd_amd.generate_sbd_data(argv)
    assert (d.method, d.init, d.seed, d.max_it, d.max_nb) == (2, 1, 12, 20, 8)
    assert (d.eps, d.carryover_type, d.ratio, d.threshold, d.max_time) == (1e-8, 3, 0.25, 0.005, 60.0)
    assert (d.do_rdm, d.bit_length, d.dump_matrix_form_wf) == (1, 64, "w.txt")
    assert (d.initial_adeterminant_bitstring, d.initial_bdeterminant_bitstring) == ("0011", "0101")
    assert sbd_amd.generate_sbd_data(["prog"]) == sbd_amd.SBD()
    with pytest.raises(ValueError):
        sbd_amd.generate_sbd_data(["prog", "--block"])
    with pytest.raises(sbd_amd.SbdError):
        sbd_amd.generate_sbd_data(["prog", "--bdet_comm_size", "2"])
