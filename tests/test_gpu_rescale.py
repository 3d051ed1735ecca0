"""GPU rescale (util.py:110-140 standard mode) is bit-identical to the oracle."""
import numpy as np
import pytest

from nixis_amd import util

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1 << 10, 400002, 3_000_001])
def test_rescale_bit_exact(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) * 1e3 + 7
    got = util.rescale(x, -9000, 9000)
    ref = oracle.rescale(x, -9000, 9000)
    assert (got == ref).all()
    assert got.min() == -9000 and got.max() == 9000


def test_rescale_constant_raises():
    with pytest.raises(ZeroDivisionError):
        util.rescale(np.full(100, 3.0), 0, 1)
    util.rescale(np.arange(5.0), 0, 1)  # the sticky error was cleared
