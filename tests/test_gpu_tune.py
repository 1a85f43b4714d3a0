"""Launch configurations and the autotuner on the GPU: every configuration
(form, tile, K slices) gives the oracle's product bit for bit, tuned choices
are cached, exported and re-imported, and tuned batched plans stay exact."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TILES = [(64, 64), (128, 64), (128, 128), (256, 256)]


def dev(g, X, p, et=np.float64):
    return g.CuModMatrix(np.asarray(X), p, elem_type=et)


def valid(form, tile, p, ksplit=1):
    L = 1
    while p > 256 ** L:
        L += 1
    if form == 3:
        return tile in ((128, 128), (256, 256)) and ksplit == 1
    if tile == (256, 256):
        return False
    if form == 1:
        return tile != (128, 64)
    return tile != (128, 128) or L <= 2


@pytest.mark.parametrize("p", [127, 65521, 8388593, 2 ** 26 - 5])
@pytest.mark.parametrize("ksplit", [1, 3, 8])
@pytest.mark.parametrize("tile", TILES, ids=str)
@pytest.mark.parametrize("form", [1, 2, 3])
def test_every_config_matches_oracle(gffm, form, tile, ksplit, p):
    if not valid(form, tile, p, ksplit):
        pytest.skip("configuration not offered for this modulus")
    g = gffm
    m, n, k = 150, 200, 2100
    print(f"gemm_ex form={form} tile={tile} ksplit={ksplit} p={p}", flush=True)
    Ah = O.draw_matrix(m, k, p, seed=21)
    Bh = O.draw_matrix(k, n, p, seed=21, offset=m * k)
    Ah[3, :] = p - 1
    Bh[:, 5] = p - 1
    C = g.zeros(np.float64, m, n, p)
    cfg = {"form": form, "tile_m": tile[0], "tile_n": tile[1], "ksplit": ksplit}
    g.mul_(C, dev(g, Ah, p), dev(g, Bh, p), config=cfg)
    assert np.array_equal(g.Array(C).astype(np.int64), O.gemm(Ah, Bh, p))


def test_k_slices_across_int32_chunks(gffm):
    # four-digit limb form, K = 70,000: three int32-safe chunks, each cut
    # into 4 concurrent slices (store, then two accumulating passes)
    g = gffm
    p, m, n, k = 2 ** 26 - 5, 40, 24, 70000
    print("gemm_ex limb ksplit=4 k=70000", flush=True)
    Ah = O.draw_matrix(m, k, p, seed=23)
    Bh = O.draw_matrix(k, n, p, seed=23, offset=m * k)
    C = g.zeros(np.float64, m, n, p)
    g.mul_(C, dev(g, Ah, p), dev(g, Bh, p), config={"form": 2, "tile_m": 64, "tile_n": 64, "ksplit": 4})
    assert np.array_equal(g.Array(C).astype(np.int64), O.gemm(Ah, Bh, p))


def test_float32_storage_with_slices(gffm):
    g = gffm
    p, m, n, k = 2039, 70, 90, 1500
    Ah = O.draw_matrix(m, k, p, seed=24)
    Bh = O.draw_matrix(k, n, p, seed=24, offset=m * k)
    for form in (1, 2):
        C = g.zeros(np.float32, m, n, p)
        g.mul_(C, dev(g, Ah, p, np.float32), dev(g, Bh, p, np.float32),
               config={"form": form, "tile_m": 64, "tile_n": 64, "ksplit": 2})
        assert np.array_equal(g.Array(C).astype(np.int64), O.gemm(Ah, Bh, p))


def test_invalid_config_rejected(gffm):
    g = gffm
    p = 2 ** 26 - 5
    A = dev(g, O.draw_matrix(8, 8, p, seed=1), p)
    C = g.zeros(np.float64, 8, 8, p)
    for cfg in ({"form": 2, "tile_m": 128, "tile_n": 128, "ksplit": 1},   # 4 digits: no 128x128
                {"form": 1, "tile_m": 128, "tile_n": 64, "ksplit": 1},    # plain: square tiles only
                {"form": 1, "tile_m": 64, "tile_n": 64, "ksplit": 0},
                {"form": 3, "tile_m": 64, "tile_n": 64, "ksplit": 1}):
        with pytest.raises(g.ArgumentError):
            g.mul_(C, A, A, config=cfg)


def test_autotune_caches_export_import(gffm):
    g = gffm
    ctx = g.Context()
    p, m, n, k = 65521, 256, 192, 1024
    cfg, ms = ctx.autotune_gemm(m, n, k, p)
    assert ms > 0
    assert cfg["form"] in (1, 2) and cfg["ksplit"] >= 1
    ctx.autotune_mode(g.TUNE_CACHED)
    got, tuned = ctx.gemm_config(m, n, k, p)
    assert tuned and got == cfg
    # same log2 bucket of every size: same entry
    got2, tuned2 = ctx.gemm_config(m - 10, n - 5, k - 100, p)
    assert tuned2 and got2 == cfg
    text = ctx.autotune_export()
    assert text.startswith("g ")
    ctx2 = g.Context()
    ctx2.autotune_mode(g.TUNE_CACHED)
    assert ctx2.gemm_config(m, n, k, p)[1] is False
    ctx2.autotune_import(text)
    assert ctx2.gemm_config(m, n, k, p) == (cfg, True)
    with pytest.raises(g.ArgumentError):
        ctx2.autotune_import("g 0 0 0 16 8 8 10 9 64 64 1 1.0\n")


def test_tune_on_miss_product_is_exact(gffm):
    g = gffm
    ctx = g.Context()
    ctx.autotune_mode(g.TUNE_ON_MISS)
    p, m, n, k = 2 ** 26 - 5, 300, 260, 1200
    Ah = O.draw_matrix(m, k, p, seed=25)
    Bh = O.draw_matrix(k, n, p, seed=25, offset=m * k)
    A = g.CuModMatrix(Ah, p, elem_type=np.float64, ctx=ctx)
    B = g.CuModMatrix(Bh, p, elem_type=np.float64, ctx=ctx)
    C = g.zeros(np.float64, m, n, p, ctx=ctx)
    g.mul_(C, A, B)
    assert ctx.gemm_config(m, n, k, p)[1]
    assert np.array_equal(g.Array(C).astype(np.int64), O.gemm(Ah, Bh, p))
    # second product of the bucket reuses the cached choice
    C2 = g.zeros(np.float64, m, n, p, ctx=ctx)
    g.mul_(C2, A, B)
    assert np.array_equal(g.Array(C2), g.Array(C))


def test_tuned_batched_plan_is_exact(gffm):
    g = gffm
    ctx = g.Context()
    ctx.autotune_mode(g.TUNE_ON_MISS)
    rng = np.random.default_rng(26)
    primes = O.largest_primes_below(2 ** 23, 8)
    A, B, C, ref = [], [], [], []
    for i in range(60):
        m, n, k = (int(x) for x in rng.integers(16, 200, size=3))
        p = int(primes[i % len(primes)])
        Ah = O.draw_matrix(m, k, p, seed=1000 + i)
        Bh = O.draw_matrix(k, n, p, seed=1000 + i, offset=m * k)
        A.append(g.CuModMatrix(Ah, p, elem_type=np.float64, ctx=ctx))
        B.append(g.CuModMatrix(Bh, p, elem_type=np.float64, ctx=ctx))
        C.append(g.zeros(np.float64, m, n, p, ctx=ctx))
        ref.append(O.gemm(Ah, Bh, p))
    jobs = [(A[i], B[i], C[i], A[i].N) for i in range(len(A))]
    g.gemm_batched(jobs)
    for Ci, Ri in zip(C, ref):
        assert np.array_equal(g.Array(Ci).astype(np.int64), Ri)
    assert "b 0 " in ctx.autotune_export()
