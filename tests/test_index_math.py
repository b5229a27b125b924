"""Index arithmetic of the HIP kernels that is exact only on a stated domain.

The x kernels find the plane k and index j of line l0 + r of a tile as
(j0 + r) / nz by a multiply-high with ceil(2^32 / nz) (tile_line_kj,
csrc/poisson_ct.h), which equals the integer quotient while (j0 + r) * nz <
2^32.  Here j0 < nz and r < 2G <= 64, and nz is an FFT length: checked for
every nz up to 8192 over every numerator the kernels can form, and shown to
fail where the bound is broken (so the check can see a wrong constant).
"""
import numpy as np


def magic(nz):
    return 0xFFFFFFFF // nz + 1  # ceil(2^32 / nz), as line_div_magic


def mulhi_div(x, nz):
    return (x.astype(np.uint64) * np.uint64(magic(nz))) >> np.uint64(32)


def test_multiply_high_division_is_exact_for_every_fft_length():
    for nz in range(1, 8193):
        x = np.arange(0, 2 * nz + 64, dtype=np.uint64)  # j0 + r < nz + 2G
        assert np.array_equal(mulhi_div(x, nz), x // np.uint64(nz)), nz


def test_multiply_high_division_fails_outside_its_domain():
    # with (j0 + r) * nz far above 2^32 the rounding of the constant shows
    nz = 3 * 2**14 + 1
    x = np.arange(2**31 - 2**20, 2**31, dtype=np.uint64)
    assert not np.array_equal(mulhi_div(x, nz), x // np.uint64(nz))
