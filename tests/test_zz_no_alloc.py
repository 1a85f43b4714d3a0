"""The allocation-free product, modp_gemm_noalloc, against modp_gemm's limb
form (same kernels, digit planes in the library's static device scratch)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

MODP_ERR_ARG = 8


def _noalloc(g, C, A, B, p):
    ctx = A.ctx
    return ctx.lib.modp_gemm_noalloc(ctx.h, A.dtype_code, g.rows(A), g.cols(B), g.cols(A), A.ptr, A.ld, B.ptr, B.ld,
                                     C.ptr, C.ld, p)


@pytest.mark.parametrize("et,p", [(np.float64, 2 ** 26 - 5), (np.float64, 8388593), (np.float64, 65521),
                                  (np.float64, 127), (np.float32, 2039)])
def test_noalloc_matches_gemm_two_calls(gffm, et, p):
    # two calls in a row, the second smaller: the scratch keeps the first
    # call's planes, and the second must not depend on them
    g = gffm
    for i, (m, n, k) in enumerate(((1500, 1300, 2100), (300, 170, 450))):
        print(f"noalloc p={p} {m}x{n}x{k}", flush=True)
        Ah = O.draw_matrix(m, k, p, seed=40 + i)
        Bh = O.draw_matrix(k, n, p, seed=40 + i, offset=m * k)
        A = g.CuModMatrix(Ah, p, elem_type=et)
        B = g.CuModMatrix(Bh, p, elem_type=et)
        C1 = g.zeros(et, m, n, p)
        C2 = g.zeros(et, m, n, p)
        g.mul_(C1, A, B, form=g.FORM_LIMB_I8)
        assert _noalloc(g, C2, A, B, p) == 0, A.ctx.lib.modp_last_error(A.ctx.h)
        c1 = g.Array(C1).astype(np.int64)
        c2 = g.Array(C2).astype(np.int64)
        assert np.array_equal(c2, c1)
        assert np.array_equal(c2, O.gemm(Ah, Bh, p))
        # the padding of C is left zero
        pad = C2.padded_array()
        assert not pad[m:, :].any() and not pad[:, n:].any()


def test_noalloc_rejects_products_beyond_the_scratch(gffm):
    # 4 digits at 2^26-5: (4096 + 4096) * 4096 * 4 bytes = 128 MiB > 64 MiB
    g = gffm
    p = 2 ** 26 - 5
    n = 4096
    A = g.rand(np.float64, n, n, p, seed=1)
    B = g.rand(np.float64, n, n, p, seed=1, offset=n * n)
    C = g.zeros(np.float64, n, n, p)
    assert _noalloc(g, C, A, B, p) == MODP_ERR_ARG
    assert b"no-allocation scratch" in A.ctx.lib.modp_last_error(A.ctx.h)
    assert not g.Array(C).any()
    # and the context still works afterwards, through both entry points
    m = 64
    Ah = O.draw_matrix(m, m, p, seed=3)
    As = g.CuModMatrix(Ah, p, elem_type=np.float64)
    Cs = g.zeros(np.float64, m, m, p)
    assert _noalloc(g, Cs, As, As, p) == 0
    assert np.array_equal(g.Array(Cs).astype(np.int64), O.gemm(Ah, Ah, p))
