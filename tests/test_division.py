"""The stepper divides the CPML psi increments by the cell size as the
product with the double nearest 1/d, rounded once to float (div_by in
fdtd2d_amd/csrc/step_kernels.hip). The identity with the correctly rounded
float division is checked here on numpy's IEEE float32 division for random
and special operands at the divisors the builders produce and at awkward
ones (powers of two, subnormals, values near the extremes)."""

from __future__ import annotations

import numpy as np
import pytest


def _operands(rng, n):
    bits = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[np.isfinite(x)]
    special = np.array([0.0, -0.0, 1.0, -1.0, np.finfo(np.float32).max, -np.finfo(np.float32).max,
                        np.finfo(np.float32).tiny, np.float32(1e-45), -np.float32(1e-45), np.inf, -np.inf],
                       np.float32)
    return np.concatenate([x, special, rng.standard_normal(n).astype(np.float32) * np.float32(1e-3)])


@pytest.mark.parametrize("d", [5e-4, 1e-3, 2.5e-4, 1e-4, 0.0009765625, 3.0, 0.1, 1e20, 1.5e-38, 7e-45, 7.77e-4])
def test_double_reciprocal_product_is_the_float_quotient(d):
    rng = np.random.default_rng(int(d * 1e6) % 2 ** 31)
    x = _operands(rng, 2_000_000)
    d32 = np.float32(d)
    with np.errstate(all="ignore"):
        want = x / d32
        got = (x.astype(np.float64) * (1.0 / np.float64(d32))).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
