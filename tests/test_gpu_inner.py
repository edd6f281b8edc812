"""The reference's InnerProduct / Normalize known answers
(tests/functionality/test_inner_product.cc:16-81) through sbd_inner, the
library's device inner product (fixed-order, compensated per-thread sums),
called on device vectors through the C ABI.  Normalize
(framework/dm_vector.h:54-87) is x / sqrt(<x, x>) with <x, x> from sbd_inner."""
import math
import sys

import numpy as np
import pytest

from conftest import has_gpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not has_gpu(), reason="needs a HIP device")]


def say(msg):
    print(f"[inner] {msg}", flush=True)
    sys.stdout.flush()


def approx_equal(a, b, tol=1e-9):
    """The reference's helper (tests/utils.h:54-66): absolute below 1e-15,
    else relative."""
    if abs(a) < 1e-15 and abs(b) < 1e-15:
        return abs(a - b) < tol
    return abs(a - b) / max(abs(a), abs(b)) < tol


@pytest.fixture(scope="module")
def ctx():
    import torch
    import sbd_amd
    from sbd_amd import synth
    torch.cuda.set_device(0)
    w = synth.make_workload("sci16", 0.25)
    m = sbd_amd.MultTPB(w["norb"], w["a"], w["b"], I0=w["I0"], h=w["h"], eri=w["eri"])
    yield m
    m.close()


def _dev(x):
    import torch
    return torch.tensor(np.asarray(x, dtype=np.float64), device="cuda")


def test_inner_product_kats(ctx):
    say("orthogonal, parallel, normalized")
    assert approx_equal(ctx.inner(_dev([1.0, 0, 0, 0]), _dev([0.0, 1, 0, 0])), 0.0)
    assert approx_equal(ctx.inner(_dev([1.0, 2, 3, 4]), _dev([2.0, 4, 6, 8])), 60.0)
    v = _dev([1.0 / math.sqrt(4.0)] * 4)
    assert approx_equal(ctx.inner(v, v), 1.0)
    say("large (N = 100000, sin x cos)")
    i = np.arange(100000, dtype=np.float64)
    X, Y = np.sin(i * 0.001), np.cos(i * 0.001)
    expected = 0.0
    for a, b in zip(X.tolist(), Y.tolist()):  # the reference's sequential sum
        expected += a * b
    assert approx_equal(ctx.inner(_dev(X), _dev(Y)), expected, 1e-10)


def test_normalize_davidson_regression(ctx):
    say("Davidson regression (dim 81796)")
    i = np.arange(81796, dtype=np.float64)
    V = _dev(np.sin(i * 0.0001) * np.exp(-i * 1e-6))
    n2 = ctx.inner(V, V)
    V = V * (1.0 / math.sqrt(n2))
    check = ctx.inner(V, V)
    assert approx_equal(check, 1.0, 1e-9)
    assert not math.isnan(check) and not math.isinf(check)
