"""Numeric identities the kernels rely on, checked on the CPU over every float
they can meet."""
import numpy as np


def test_forward_dct_rounding_in_fp32():
    """mb_kernel rounds each postscaled forward-DCT coefficient x as
    floorf(x + 0.49999f).  The reference writes floorf(x + 0.49999)
    (src/mbgpu_kernels.cuh:413-437): float plus a double literal is a double
    sum, which floorf's float parameter takes rounded to float.  The two agree
    for every float in the binades 2^-10 .. 2^15 (all mantissas, both signs).
    Below 2^-10 both sums lie within (0.498, 0.502) and floor to 0; the
    coefficients stay below 2^13."""
    half = np.float32(0.49999)
    mant = np.arange(1 << 23, dtype=np.uint32)
    for e in range(-10, 15):
        for sign in (0, 1 << 31):
            x = (mant | np.uint32(((e + 127) << 23) | sign)).view(np.float32)
            ref = np.floor((x.astype(np.float64) + 0.49999).astype(np.float32))
            got = np.floor(x + half)
            bad = np.flatnonzero(ref != got)
            assert bad.size == 0, f"binade 2^{e}: x = {x[bad[0]]!r}"
    tiny = np.float32(2.0 ** -10)
    for x in (tiny, -tiny, np.float32(0.0)):
        assert np.floor(np.float32(x) + half) == 0.0
        assert np.floor(np.float32(np.float64(x) + 0.49999)) == 0.0


def test_forward_dct_rounding_fp32_catches_a_wrong_constant():
    """The check above is sharp: the float sum with 0.5f in place of
    0.49999f differs at the half-integers."""
    x = np.float32([0.5, 1.5, -0.5, 100.5])
    ref = np.floor((x.astype(np.float64) + 0.49999).astype(np.float32))
    assert not np.array_equal(np.floor(x + np.float32(0.5)), ref)
