"""Batched products in both arithmetic forms: the int8 digit form (default for
k <= 16384, one class per digit count 1..4, with one 64-row panel per work
item for k <= 256 and 64 x 64 tiles otherwise) and the FP64 form (forced with
MODP_BATCHED_FP64=1, and used for deeper contractions), each product equal to
the oracle's."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def make_jobs(g, specs, fill=None, seed=0):
    jobs, host = [], []
    for i, (m, n, k, p) in enumerate(specs):
        if fill is None:
            Ah = O.draw_matrix(m, k, p, seed=seed * 1000 + i)
            Bh = O.draw_matrix(k, n, p, seed=seed * 1000 + i, offset=m * k)
        else:
            Ah = np.full((m, k), fill(p), dtype=np.int64)
            Bh = np.full((k, n), fill(p), dtype=np.int64)
        jobs.append((g.CuModMatrix(Ah, p, elem_type=np.float64), g.CuModMatrix(Bh, p, elem_type=np.float64),
                     g.zeros(np.float64, m, n, p), p))
        host.append((Ah, Bh, p))
    return jobs, host


def check(g, jobs, host):
    g.gemm_batched(jobs)
    g.context().synchronize()
    for i, ((_, _, C, _), (Ah, Bh, p)) in enumerate(zip(jobs, host)):
        assert np.array_equal(g.Array(C).astype(np.int64), O.gemm(Ah, Bh, p)), f"product {i} (p={p})"


P4 = O.largest_primes_below(2 ** 25, 1)[0]
PRIMES = [2, 3, 251, 257, 65521, 65537, 8388593, P4, 2 ** 26 - 5]


@pytest.mark.parametrize("p", PRIMES)
def test_digit_form_each_digit_count(gffm, p):
    rng = np.random.default_rng(p % 1000)
    specs = [(int(rng.integers(1, 200)), int(rng.integers(1, 200)), int(rng.integers(1, 300)), p) for _ in range(12)]
    specs += [(64, 64, 64, p), (65, 63, 129, p), (1, 1, 1, p), (128, 7, 2000, p)]
    print(f"digit form p={p}", flush=True)
    check(gffm, *make_jobs(gffm, specs, seed=5))


@pytest.mark.parametrize("fill", ["pm1", "half"])
@pytest.mark.parametrize("p", [65521, 8388593, 2 ** 26 - 5])
def test_digit_form_extreme_residues(gffm, p, fill):
    f = (lambda q: q - 1) if fill == "pm1" else (lambda q: (q - 1) // 2)
    specs = [(70, 66, 3000, p), (33, 40, 256, p), (64, 64, 16384, p)]
    print(f"digit form extremes p={p} fill={fill}", flush=True)
    check(gffm, *make_jobs(gffm, specs, fill=f))


@pytest.mark.parametrize("tiles", ["0", "1"])
def test_digit_form_row_panels(gffm, monkeypatch, tiles):
    # k <= 256: one item per 64-row panel, whose A digits stay in LDS over all
    # column tiles; MODP_BATCHED_TILES=1 puts the same products on 64 x 64 tiles
    monkeypatch.setenv("MODP_BATCHED_TILES", tiles)
    specs = []
    for p in [3, 251, 65521, 8388593, 2 ** 26 - 5]:
        specs += [(300, 500, 256, p), (129, 65, 1, p), (64, 1000, 33, p), (200, 200, 257, p), (7, 130, 255, p)]
    print(f"row panels tiles={tiles}", flush=True)
    check(gffm, *make_jobs(gffm, specs, seed=13))
    fill = [(lambda q: q - 1), (lambda q: (q - 1) // 2)]
    for f in fill:
        check(gffm, *make_jobs(gffm, [(130, 300, 256, 2 ** 26 - 5), (65, 66, 256, 65521)], fill=f))


def test_fp64_form_forced_vs_oracle(gffm, monkeypatch):
    monkeypatch.setenv("MODP_BATCHED_FP64", "1")
    rng = np.random.default_rng(3)
    primes = O.largest_primes_below(2 ** 23, 64)
    specs = [(int(rng.integers(16, 257)), int(rng.integers(16, 257)), int(rng.integers(16, 257)),
              primes[int(rng.integers(0, 64))]) for _ in range(120)]
    specs += [(50, 40, 600, 2 ** 26 - 5), (3, 5, 7, 65521)]
    check(gffm, *make_jobs(gffm, specs, seed=7))


def test_deep_contraction_takes_fp64_form(gffm):
    # k above the digit form's 16384 goes to the FP64 classes, in the same plan
    specs = [(20, 18, 17000, 65521), (9, 30, 16385, 2 ** 26 - 5), (40, 40, 100, 8388593)]
    check(gffm, *make_jobs(gffm, specs, seed=11))


@pytest.mark.parametrize("fp64", [False, True])
def test_unreduced_inputs_are_reduced(gffm, monkeypatch, fp64):
    # stored values outside [0, p) (negative, >= p, up to 2^40) are reduced
    # first, as in the single-product path
    if fp64:
        monkeypatch.setenv("MODP_BATCHED_FP64", "1")
    g = gffm
    rng = np.random.default_rng(17)
    jobs, host = [], []
    for i, p in enumerate([101, 65521, 8388593, 2 ** 26 - 5]):
        m, n, k = 37 + i, 29 + 3 * i, 70 + 11 * i
        Ah = rng.integers(-(2 ** 40), 2 ** 40, size=(m, k))
        Bh = rng.integers(-3 * p, 3 * p, size=(k, n))
        Ah[:, ::3] = np.mod(Ah[:, ::3], p)  # some columns in range
        A = g.CuModMatrix(Ah, p, mod=False, elem_type=np.float64)
        B = g.CuModMatrix(Bh, p, mod=False, elem_type=np.float64)
        jobs.append((A, B, g.zeros(np.float64, m, n, p), p))
        host.append((np.mod(Ah, p), np.mod(Bh, p), p))
    check(g, jobs, host)


@pytest.mark.parametrize("fp64", [False, True])
def test_float32_storage(gffm, monkeypatch, fp64):
    # residues stored as float32; the reference's Float32 rule admits only
    # moduli whose products fit the 24-bit significand (stripe_mul.jl:200-202)
    if fp64:
        monkeypatch.setenv("MODP_BATCHED_FP64", "1")
    g = gffm
    jobs, host = [], []
    for i, p in enumerate([2, 7, 251, O.largest_primes_below(2898, 1)[0]]):
        m, n, k = 40 + 7 * i, 33 + 5 * i, 90 + 13 * i
        Ah = O.draw_matrix(m, k, p, seed=60 + i)
        Bh = O.draw_matrix(k, n, p, seed=60 + i, offset=m * k)
        jobs.append((g.CuModMatrix(Ah, p, elem_type=np.float32), g.CuModMatrix(Bh, p, elem_type=np.float32),
                     g.zeros(np.float32, m, n, p), p))
        host.append((Ah, Bh, p))
    g.gemm_batched(jobs, elem_type=np.float32)
    for i, ((_, _, C, _), (Ah, Bh, p)) in enumerate(zip(jobs, host)):
        assert np.array_equal(g.Array(C).astype(np.int64), O.gemm(Ah, Bh, p)), f"product {i} (p={p})"
