"""The community mix of the hub and giant kernels is a bijection of the
32-bit ids (its inverse recovers every input), and it spreads consecutive
ids evenly over the partitions its top bits select."""
import numpy as np

from partition_mix import fmix32, fmix32_inverse


def test_mix_inverse_recovers_every_input():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.integers(0, 1 << 32, 1 << 20, dtype=np.uint64),
                        np.arange(4096, dtype=np.uint64),
                        np.array([0, 1, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFE, 0xFFFFFFFF],
                                 np.uint64)])
    assert np.array_equal(fmix32_inverse(fmix32(x)), x)
    assert np.array_equal(fmix32(fmix32_inverse(x)), x)


def test_consecutive_ids_spread_over_partitions():
    ids = np.arange(1 << 20, dtype=np.uint64)
    for bits in (5, 8, 12):
        part = fmix32(ids) >> (32 - bits)
        cnt = np.bincount(part.astype(np.int64), minlength=1 << bits)
        mean = len(ids) / (1 << bits)
        tol = 6 * np.sqrt(mean)   # six binomial standard deviations
        assert cnt.min() > mean - tol and cnt.max() < mean + tol, (bits, cnt.min(), cnt.max())
