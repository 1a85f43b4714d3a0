"""The multi-modular (RNS) int8 form of the mod-p GEMM against the oracle:
shapes, moduli on both sides of each modulus-count step, the extreme
residues that reach the reconstruction bound, K chunks, element types, the
single-plane form of moduli up to 255 and the elimination's subtracting
epilogue (through the autotuner's configs)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

RNS = 3


def dev(g, X, p, et=np.float64):
    return g.CuModMatrix(np.asarray(X), p, elem_type=et)


def rns_product(g, Ah, Bh, p, et=np.float64):
    m, k = Ah.shape
    n = Bh.shape[1]
    C = g.zeros(et, m, n, p)
    g.mul_(C, dev(g, Ah, p, et), dev(g, Bh, p, et), form=RNS)
    return g.Array(C).astype(np.int64)


@pytest.mark.parametrize("p", [2, 3, 127, 251, 65521, 8388593, 2 ** 24 - 3, 2 ** 26 - 5, 2 ** 26])
@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 5, 3), (130, 257, 129), (300, 200, 1000), (513, 260, 2100)],
                         ids=str)
def test_rns_vs_oracle(gffm, p, shape):
    g = gffm
    m, n, k = shape
    print(f"rns p={p} shape={shape}", flush=True)
    Ah = O.draw_matrix(m, k, p, seed=21)
    Bh = O.draw_matrix(k, n, p, seed=21, offset=m * k)
    assert np.array_equal(rns_product(g, Ah, Bh, p), O.gemm(Ah, Bh, p))


@pytest.mark.parametrize("p", [65521, 8388593, 2 ** 26 - 5, 2 ** 26 - 1, 2 ** 26])
@pytest.mark.parametrize("fill", ["half", "half_plus", "pm1", "mixed"])
def test_rns_extreme_residues(gffm, p, fill):
    # |a b| = floor(p/2)^2 in every product: the largest integer the
    # reconstruction must recover (all of one sign), and mixed signs
    g = gffm
    h = p // 2
    m, n, k = 70, 90, 3000
    val = {"half": h, "half_plus": h + 1 if h + 1 < p else h, "pm1": p - 1}.get(fill)
    if fill == "mixed":
        rng = np.random.default_rng(p)
        Ah = rng.choice([h, (h + 1) % p, 0, p - 1], size=(m, k))
        Bh = rng.choice([h, (h + 1) % p, 1, p - 1], size=(k, n))
    else:
        Ah = np.full((m, k), val, dtype=np.int64)
        Bh = np.full((k, n), val, dtype=np.int64)
    assert np.array_equal(rns_product(g, Ah, Bh, p), O.gemm(Ah, Bh, p))


def test_rns_k_chunks(gffm):
    # K above one int32-exact chunk (130944): two chunks, the second added
    # mod p by the combine step; extreme residues in the first rows
    g = gffm
    p = 2 ** 26 - 5
    m, n, k = 40, 33, 140000
    Ah = O.draw_matrix(m, k, p, seed=22)
    Bh = O.draw_matrix(k, n, p, seed=22, offset=m * k)
    Ah[0, :] = p // 2
    Bh[:, 0] = p // 2
    assert np.array_equal(rns_product(g, Ah, Bh, p), O.gemm(Ah, Bh, p))


def test_rns_float32_storage(gffm):
    g = gffm
    p = 2039
    Ah = O.draw_matrix(300, 5000, p, seed=23)
    Bh = O.draw_matrix(5000, 200, p, seed=23, offset=300 * 5000)
    assert np.array_equal(rns_product(g, Ah, Bh, p, np.float32), O.gemm(Ah, Bh, p))


@pytest.mark.parametrize("tile", [128, 256])
def test_rns_explicit_configs(gffm, tile):
    # both output tiles through modp_gemm_ex, at a size where 256 pads
    g = gffm
    ctx = g.context()
    p = 2 ** 26 - 5
    m, n, k = 700, 450, 1500
    Ah = O.draw_matrix(m, k, p, seed=24)
    Bh = O.draw_matrix(k, n, p, seed=24, offset=m * k)
    A, B = dev(g, Ah, p), dev(g, Bh, p)
    C = g.zeros(np.float64, m, n, p)
    cfg = g._lib.GemmConfig(RNS, tile, tile, 1)
    import ctypes
    ctx.check(ctx.lib.modp_gemm_ex(ctx.h, ctypes.byref(cfg), 0, m, n, k, A.ptr, A.ld, B.ptr, B.ld, C.ptr, C.ld, p))
    assert np.array_equal(g.Array(C).astype(np.int64), O.gemm(Ah, Bh, p))
    # K slices are not a configuration of this form
    bad = g._lib.GemmConfig(RNS, tile, tile, 2)
    assert ctx.lib.modp_gemm_ex(ctx.h, ctypes.byref(bad), 0, m, n, k, A.ptr, A.ld, B.ptr, B.ld, C.ptr, C.ld, p) != 0


@pytest.mark.parametrize("p", [2, 3, 4, 9, 127, 128, 200, 251, 254, 255])
@pytest.mark.parametrize("fill", ["half", "half_plus", "pm1", "mixed"])
def test_rns_single_plane_extremes(gffm, p, fill):
    # 2 <= p <= 255: one int8 product modulo p itself (balanced residues up to
    # 127 in magnitude, ties of even moduli at +-p/2), composite moduli too
    g = gffm
    h = p // 2
    m, n, k = 66, 70, 3000
    val = {"half": h, "half_plus": (h + 1) % p, "pm1": p - 1}.get(fill)
    if fill == "mixed":
        rng = np.random.default_rng(p + 7)
        Ah = rng.choice([h, (h + 1) % p, 0, p - 1], size=(m, k))
        Bh = rng.choice([h, (h + 1) % p, 1, p - 1], size=(k, n))
    else:
        Ah = np.full((m, k), val, dtype=np.int64)
        Bh = np.full((k, n), val, dtype=np.int64)
    assert np.array_equal(rns_product(g, Ah, Bh, p), O.gemm(Ah, Bh, p))


@pytest.mark.parametrize("p", [251, 255])
def test_rns_single_plane_k_chunks(gffm, p):
    # two int32-exact chunks (130944 + rest) at the largest balanced
    # magnitude, 127^2 per product in the first chunk's int32 sums
    g = gffm
    m, n, k = 5, 6, 131000
    Ah = O.draw_matrix(m, k, p, seed=26)
    Bh = O.draw_matrix(k, n, p, seed=26, offset=m * k)
    Ah[0, :] = p // 2
    Bh[:, 0] = p // 2
    Ah[1, :] = (p + 1) // 2
    Bh[:, 1] = (p + 1) // 2
    assert np.array_equal(rns_product(g, Ah, Bh, p), O.gemm(Ah, Bh, p))


def test_rns_single_plane_float32(gffm):
    g = gffm
    p = 251
    Ah = O.draw_matrix(300, 5000, p, seed=27)
    Bh = O.draw_matrix(5000, 200, p, seed=27, offset=300 * 5000)
    assert np.array_equal(rns_product(g, Ah, Bh, p, np.float32), O.gemm(Ah, Bh, p))


def test_auto_picks_rns_for_large_moduli(gffm):
    g = gffm
    ctx = g.context()
    cfg, tuned = ctx.gemm_config(8192, 8192, 8192, 2 ** 26 - 5)
    assert cfg["form"] == RNS and cfg["tile_m"] == 256
    cfg, _ = ctx.gemm_config(8192, 8192, 8192, 65521)
    assert cfg["form"] == 2
    # moduli up to 255: one plane, chosen over the one-digit limb form while
    # its grid fills the chip
    cfg, _ = ctx.gemm_config(8192, 8192, 8192, 127)
    assert cfg["form"] == RNS and cfg["tile_m"] == 256
    for n in (1024, 4096):
        cfg, _ = ctx.gemm_config(n, n, n, 127)
        assert cfg["form"] == 2
