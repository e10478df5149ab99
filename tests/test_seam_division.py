"""The seam pass finds the block column of a pixel by a float-reciprocal
division with one correction each way (div_small in nf3d_run.h).  This checks,
in float32 arithmetic with the GPU's rounding (round-to-nearest products,
truncating conversion), that the method equals integer division for every
column offset of planes up to 2^17 pixels wide and every step 1..64, and on a
random sample up to its stated limit 2^22; the last test shows the
corrections repair a reciprocal one ulp off."""
import numpy as np


def div_small(n, d, inv=None):
    if inv is None:
        inv = np.float32(1.0) / np.float32(d)
    q = (n.astype(np.float32) * inv).astype(np.int32)  # truncation toward zero
    r = n - q * d
    q = q + (r >= d)
    q = q - (r < 0)
    return q


def test_exhaustive_small_planes():
    n = np.arange(1 << 17, dtype=np.int32)
    for d in range(1, 65):
        assert np.array_equal(div_small(n, d), n // d), d


def test_sample_to_limit():
    rng = np.random.default_rng(7)
    n = rng.integers(0, 1 << 22, size=1 << 20, dtype=np.int32)
    for d in range(1, 65):
        assert np.array_equal(div_small(n, d), n // d), d


def test_corrections_cover_a_reciprocal_one_ulp_off():
    # with 1/d one ulp low the bare quotient falls one short at exact
    # multiples; the corrected one is exact with 1/d one ulp low or high
    n = np.arange(1 << 20, dtype=np.int32)
    for d in (3, 5, 7, 48, 49, 63):
        inv = np.float32(1.0) / np.float32(d)
        low, high = np.nextafter(inv, np.float32(0)), np.nextafter(inv, np.float32(1))
        bare = (n.astype(np.float32) * low).astype(np.int32)
        assert np.count_nonzero(bare != n // d) > 0, d
        for off in (low, high):
            assert np.array_equal(div_small(n, d, off), n // d), d
