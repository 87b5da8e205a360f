"""The matrix-core accumulation order: v_mfma_f32_16x16x4f32 computes
fma(a3, b3, fma(a2, b2, fma(a1, b1, fma(a0, b0, c)))) — four correctly rounded
fused multiply-adds in ascending k — which is the order the oracle restates for
the surrogate network (z = fma chain over the padded inputs).  Crafted cases
separate the candidates (fused vs rounded products, k order vs pairwise vs one
rounding of the exact sum, ties, cancellation, denormals); random cases check
the chain bitwise."""
from fractions import Fraction as Fr

import numpy as np
import pytest

from sce_amd import gpu as G

pytestmark = pytest.mark.gpu


def _r32(x):
    return np.float32(float(x))


def _fma_chain(a, b, c):
    acc = np.float32(c)
    for k in range(4):
        acc = _r32(Fr(float(np.float32(a[k]))) * Fr(float(np.float32(b[k]))) + Fr(float(acc)))
    return acc


def _exact_once(a, b, c):
    return _r32(sum(Fr(float(np.float32(a[k]))) * Fr(float(np.float32(b[k]))) for k in range(4))
                + Fr(float(np.float32(c))))


E, U = 2.0 ** 24, 1.0 / 4096
CRAFTED = [
    (1.0, [U, U, 0, 0], [U, U, 0, 0]),
    (-1.0, [1 + U / 2, 0, 0, 0], [1 + U / 2, 0, 0, 0]),
    (0.0, [E, 1, 1, 1], [1, 1, 1, 1]),
    (E, [1, 1, 1, 0], [1, 1, 1, 0]),
    (0.0, [1, 1, 1, E], [1, 1, 1, 1]),
    (1.0, [E, -E, 0, 0], [1, 1, 0, 0]),
    (0.0, [1e-30, 1e-30, 0, 0], [1e-10, 1e-10, 0, 0]),
]


def test_crafted_cases_follow_the_fma_chain():
    a = np.array([t[1] for t in CRAFTED], np.float32)
    b = np.array([t[2] for t in CRAFTED], np.float32)
    c = np.array([t[0] for t in CRAFTED], np.float32)
    got = G.dev_mfma(a, b, c)
    for k in range(len(c)):
        assert got[k].view(np.uint32) == _fma_chain(a[k], b[k], c[k]).view(np.uint32), k
    # the cases do separate the chain from one rounding of the exact sum
    assert any(_fma_chain(a[k], b[k], c[k]) != _exact_once(a[k], b[k], c[k]) for k in range(len(c)))


def test_random_cases_follow_the_fma_chain():
    rng = np.random.default_rng(11)
    n = 4000
    a = (rng.standard_normal((n, 4)) * np.exp2(rng.integers(-20, 20, (n, 4)))).astype(np.float32)
    b = (rng.standard_normal((n, 4)) * np.exp2(rng.integers(-20, 20, (n, 4)))).astype(np.float32)
    c = (rng.standard_normal(n) * np.exp2(rng.integers(-20, 20, n))).astype(np.float32)
    got = G.dev_mfma(a, b, c)
    want = np.array([_fma_chain(a[k], b[k], c[k]) for k in range(n)], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
