"""The C++ host layer (include/sbd_tpb.hpp: FCIDump/LoadFCIDump, LoadAlphaDets,
from_string, sort_bitarray, tpb::SBD, tpb::diag in both forms, the MultTPB
operator plug-in and Davidson on it) through tests/cpp/test_tpb_host, a test
program in the style of the reference's own C++ tests.  Host-only checks run
here; the solves on H2O 1e-3 (the reference-held energy, density sums, RDM
trace, save and warm start, plug-in Davidson, accumulating run) run on the GPU."""
import gzip
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT, has_gpu

BIN = os.path.join(ROOT, "tests", "cpp", "test_tpb_host")


def _binary():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)
    return BIN


def _inputs(tmp_path):
    out = []
    for name in ("fcidump.txt.gz", "h2o-1em3-alpha.txt.gz"):
        dst = tmp_path / name[:-3]
        with gzip.open(os.path.join(GOLDEN, "h2o", name), "rb") as f, open(dst, "wb") as g:
            shutil.copyfileobj(f, g)
        out.append(str(dst))
    return out


def test_cpp_host_layer_host_only(tmp_path):
    fd, alpha = _inputs(tmp_path)
    r = subprocess.run([_binary(), "--host", fd, alpha], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout


@pytest.mark.gpu
@pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")
def test_cpp_host_layer_diag_h2o(tmp_path):
    fd, alpha = _inputs(tmp_path)
    print("[gpu] C++ host layer: diag h2o 1e-3", flush=True)
    r = subprocess.run([_binary(), "--gpu", fd, alpha], cwd=tmp_path, capture_output=True,
                       text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout
