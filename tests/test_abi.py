"""The C ABI library loads and exports every entry point include/modp.h
declares; host-side logic that needs no GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "modp.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(modp_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["modp_gemm", "modp_pluq", "modp_lu", "modp_rref", "modp_gemm_batched", "modp_elim_batched",
                 "modp_apply_row_perm", "modp_apply_col_perm", "modp_ctx_create", "modp_mod"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from gpufinitefieldmatrices import _lib
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []
    # the Python binding declares a signature for every symbol
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(declared_functions()) == bound


def test_version_and_no_device_here():
    from gpufinitefieldmatrices import _lib
    lib = _lib.load()
    assert lib.modp_version().startswith(b"modp_hip")
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and not os.path.exists("/dev/kfd"):
        n = ctypes.c_int(-1)
        st = lib.modp_device_count(ctypes.byref(n))
        assert st in (0, _lib.MODP_ERR_NO_DEVICE) and n.value == 0
        h = ctypes.c_void_p()
        assert lib.modp_ctx_create(0, ctypes.byref(h)) == _lib.MODP_ERR_NO_DEVICE


def test_null_context_is_rejected():
    from gpufinitefieldmatrices import _lib
    lib = _lib.load()
    assert lib.modp_synchronize(None) == _lib.MODP_ERR_ARG
    assert lib.modp_gemm(None, 0, 0, 1, 1, 1, None, 1, None, 1, None, 1, 7) == _lib.MODP_ERR_ARG


def test_find_max_stripe_ops_matches_reference_table():
    import gpufinitefieldmatrices as g
    # stripe_mul.jl:13-27 values (closed form floor((2^bits-1)/(N-1)^2) - 1)
    assert g.find_max_stripe_ops(np.float64, 65521) == 2098175
    assert g.find_max_stripe_ops(np.float64, 2 ** 26 - 5) == 1
    assert g.find_max_stripe_ops(np.float64, 127) == 567346891832
    assert g.find_max_stripe_ops(np.float64, 2 ** 26) == 1
    assert g.find_max_stripe_ops(np.float64, 2 ** 26 + 1) == 0
    assert g.find_max_stripe_ops(np.float32, 127) == 1055
    assert g.find_max_stripe_ops(np.float32, 2 ** 11) == 3
    assert g.find_max_stripe_ops(np.float32, 11 ** 3) == 8
    assert g.find_max_stripe_ops(np.float32, 2897) == 1
    assert g.find_max_stripe_ops(np.float32, 2898) == 0
    assert g.find_max_stripe_ops(np.float32, 65521) == -1
    with pytest.raises(g.ArgumentError):
        g.find_max_stripe_ops(np.int64, 7)


def test_mod_inv_host_mirror():
    import gpufinitefieldmatrices as g
    for p in (7, 65521):
        for a in (1, 2, 3, p - 1):
            assert (a * g.mod_inv(a, p)) % p == 1


def test_constructor_rejects_huge_modulus_before_device_use():
    import gpufinitefieldmatrices as g
    with pytest.raises(g.CuModArrayModulusMismatchException):
        g.CuModMatrix([[1]], 2 ** 52 + 1)
