"""The FP32 residue reductions of the multi-modular GEMM, checked for every
input they can receive.

The splits (gemm_rns.h, res8/pack16q) reduce an exact FP32 integer t with
r = t - q * rint(t * fl(1/q)); the core's epilogue reduces t = a (2^16 mod q)
+ b the same way and adds q to negative results.  numpy's float32 multiply
and rint round like the device's v_mul_f32 and v_rndne_f32, and the fma of
the device is exact here (every term is an integer below 2^24), so this
evaluates the device formula bit for bit over the whole input range.
"""
import numpy as np
import pytest

QLIST = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 211]


def _reduce(t, q):
    tf = t.astype(np.float32)
    assert np.array_equal(tf.astype(np.int64), t)  # t is exact in FP32
    iq = np.float32(1.0) / np.float32(q)
    n = np.rint(tf * iq).astype(np.int64)
    return t - n * q


def _chunks(lo, hi, step=1 << 22):
    for s in range(lo, hi + 1, step):
        yield np.arange(s, min(s + step, hi + 1), dtype=np.int64)


@pytest.mark.parametrize("q", QLIST)
def test_split_residue_fits_int8(q):
    # A side: |t| <= 2^13 (q - 1) + 4095; B side: t = xh (4096 y mod q) + xl y
    # with y < q: -2^13 (q - 1) <= t <= 2^13 (q - 1) + 4095 (q - 1)
    lo, hi = -(2 ** 13) * (q - 1), 2 ** 13 * (q - 1) + 4095 * (q - 1)
    for t in _chunks(lo, hi):
        r = _reduce(t, q)
        if q == 256:
            assert r.min() >= -128 and r.max() <= 128
        else:
            assert np.abs(r).max() <= (q - 1) // 2  # the balanced residue
        b = r & 255
        b8 = np.where(b >= 128, b - 256, b)  # the stored int8
        assert np.all((b8 - t) % q == 0)


@pytest.mark.parametrize("q", QLIST)
def test_epilogue_reduction_is_canonical(q):
    c16 = 65536 % q
    lo, hi = -(2 ** 15) * c16, (2 ** 15 - 1) * c16 + 65535
    for t in _chunks(lo, hi):
        r = _reduce(t, q)
        e = np.where(r < 0, r + q, r)
        assert np.array_equal(e, t % q)


def test_epilogue_split_of_int32_is_exact():
    # v = 2^16 a + b with a = v >> 16 (arithmetic) and b = v & 0xFFFF
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.integers(-2 ** 31, 2 ** 31, 100000, dtype=np.int64),
                        np.array([-2 ** 31, 2 ** 31 - 1, -1, 0, 65535, -65536])])
    a, b = v >> 16, v & 0xFFFF
    assert np.array_equal(a * 65536 + b, v)
    for q in QLIST:
        t = a * (65536 % q) + b
        assert np.array_equal(t % q, v % q)
