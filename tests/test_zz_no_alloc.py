"""sbd_inner_noalloc (the inner product with its reduction scratch in a fixed
device array) against sbd_inner, both called through the C ABI: the same
bits on consecutive calls whose sizes shrink (the second call reuses the
array the first one filled), on 16-byte-aligned and unaligned vectors, and
on an empty range; both agree with a compensated host sum."""
import ctypes as C
import math
import sys

import numpy as np
import pytest

from conftest import has_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]


def say(msg):
    print(f"[noalloc] {msg}", flush=True)
    sys.stdout.flush()


def _call(fn, ctx, px, py, n):
    out = C.c_double(float("nan"))
    rc = fn(ctx, C.c_void_p(px), C.c_void_p(py), C.c_uint64(n), C.byref(out), None)
    return rc, out.value


def test_inner_noalloc_matches_inner():
    import torch
    import sbd_amd
    from sbd_amd import synth
    torch.cuda.set_device(0)
    L = sbd_amd.lib()
    w = synth.make_workload("sci16", 0.5)
    m = sbd_amd.MultTPB(w["norb"], w["a"], w["b"], I0=w["I0"], h=w["h"], eri=w["eri"])
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    big = torch.randn(3_000_001, device="cuda", dtype=torch.float64, generator=g)
    other = torch.randn(3_000_001, device="cuda", dtype=torch.float64, generator=g)
    torch.cuda.synchronize()
    # (offset, n): a large call, then smaller ones; odd offsets are 8 bytes
    # off the 16-byte alignment and take the element-by-element loads
    cases = [(0, 3_000_000), (0, 70_001), (1, 1_000_000), (1, 777), (0, 1), (3, 0)]
    for off, n in cases:
        say(f"offset {off} n {n}")
        x, y = big[off:off + n], other[off:off + n]
        # element addresses from the base pointers (a zero-length view has none)
        px, py = big.data_ptr() + 8 * off, other.data_ptr() + 8 * off
        rc_a, a = _call(L.sbd_inner_noalloc, m._ctx, px, py, n)
        rc_b, b = _call(L.sbd_inner, m._ctx, px, py, n)
        assert rc_a == 0 and rc_b == 0, (rc_a, rc_b)
        assert a == b, (off, n, a, b)
        X, Y = x.cpu().numpy(), y.cpu().numpy()
        ref = math.fsum((X * Y).tolist()) if n else 0.0
        assert abs(a - ref) <= 1e-13 * max(1.0, float(np.abs(X * Y).sum())), (a, ref)
    # argument checks, before any device work
    out = C.c_double(0.0)
    assert L.sbd_inner_noalloc(m._ctx, None, C.c_void_p(other.data_ptr()), C.c_uint64(4),
                               C.byref(out), None) == 1
    m.close()
