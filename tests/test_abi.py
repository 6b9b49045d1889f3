"""CPU-side checks of the C ABI: the library loads without a GPU and exports
every function include/alto_hip.h declares; host-side argument validation."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "alto_hip.h")
LIB = os.path.join(ROOT, "alto_amd", "libalto_hip.so")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(alto_[a-z0-9_]+)\s*\(", src)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="libalto_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libalto_hip.so not built")
def test_invalid_arguments_rejected_without_gpu_work():
    import alto_amd as A

    L = A.lib()
    out = ctypes.c_void_p()
    dims = (ctypes.c_uint64 * 3)(10, 10, 10)
    # nmodes out of range, null pointers, bad lit width: rejected on the host
    assert L.alto_tensor_create(0, dims, 5, None, None, 0, 0, 0, 1, None, ctypes.byref(out)) == A.ALTO_OK - 1
    assert L.alto_tensor_create(6, dims, 5, None, None, 0, 0, 0, 1, None, ctypes.byref(out)) < 0
    assert b"invalid" in L.alto_last_error()
    assert L.alto_tensor_destroy(None) == 0
    assert L.alto_mttkrp(None, 0, 16, None, None, 0, 0, None) < 0
    # multi-tensor MTTKRP: negative count, null tables, null tensor, unknown
    # precision; an empty batch is a no-op
    assert L.alto_mttkrp_batched(-1, None, None, 16, None, None, 0, 0, None) < 0
    assert L.alto_mttkrp_batched(2, None, None, 16, None, None, 0, 0, None) < 0
    one = (ctypes.c_void_p * 1)(None)
    modes = (ctypes.c_int * 1)(0)
    assert L.alto_mttkrp_batched(1, one, modes, 16, one, one, 0, 0, None) < 0
    assert b"null" in L.alto_last_error()
    assert L.alto_mttkrp_batched(0, None, None, 16, None, None, 9, 0, None) < 0
    assert L.alto_mttkrp_batched(0, None, None, 16, None, None, 0, 0, None) == A.ALTO_OK
    # lockstep sessions: negative counts, null table, null session; none is a no-op
    assert L.alto_cpd_sessions_iterate(-1, None, 1, 0.0, None, None) < 0
    assert L.alto_cpd_sessions_iterate(1, None, 1, 0.0, None, None) < 0
    assert L.alto_cpd_sessions_iterate(1, one, 1, 0.0, None, None) < 0
    assert L.alto_cpd_sessions_iterate(1, one, -1, 0.0, None, None) < 0
    assert L.alto_cpd_sessions_iterate(0, None, 5, 0.0, None, None) == A.ALTO_OK
    assert L.alto_version().startswith(b"alto-mi355x")


@pytest.mark.skipif(not os.path.exists(LIB), reason="libalto_hip.so not built")
def test_dropin_entry_points_reject_bad_arguments_on_the_host():
    """The entry points behind the C++ drop-in (COO MTTKRP, Kruskal
    normalisation, fit from parts, CP-APR with a stop rule) reject invalid
    arguments before any device work; a host coordinate out of range is
    found by the range check."""
    import numpy as np

    import alto_amd as A

    L = A.lib()
    dims = (ctypes.c_uint64 * 3)(4, 5, 6)
    U = [np.zeros((d, 2)) for d in (4, 5, 6)]
    up = (ctypes.c_void_p * 3)(*[u.ctypes.data for u in U])
    cidx = np.array([[0, 3], [1, 4], [2, 6]], np.uint64)  # 6 is out of range in mode 2
    cp = (ctypes.c_void_p * 3)(*[cidx[m].ctypes.data for m in range(3)])
    vals = np.ones(2)
    out = np.zeros((4, 2))
    assert L.alto_coo_mttkrp(0, dims, 2, cp, vals.ctypes.data, 0, 2, up, out.ctypes.data, 1, 0, None) < 0
    assert L.alto_coo_mttkrp(3, dims, 2, cp, vals.ctypes.data, 3, 2, up, out.ctypes.data, 1, 0, None) < 0
    assert L.alto_coo_mttkrp(3, dims, 2, cp, vals.ctypes.data, 0, 0, up, out.ctypes.data, 1, 0, None) < 0
    assert L.alto_coo_mttkrp(3, dims, 2, cp, vals.ctypes.data, 0, 2, up, out.ctypes.data, 1, 7, None) < 0
    assert L.alto_coo_mttkrp(3, dims, 2, cp, vals.ctypes.data, 0, 2, up, out.ctypes.data, 1, 0, None) < 0
    assert b"out of range" in L.alto_last_error()
    lam = np.ones(2)
    assert L.alto_kruskal_normalize(3, dims, 0, up, lam.ctypes.data, 0, 0, 0, None) < 0
    assert L.alto_kruskal_normalize(3, dims, 2, up, lam.ctypes.data, -1, A.KRUSKAL_NORM2, 0, None) < 0
    assert L.alto_kruskal_normalize(3, dims, 2, up, lam.ctypes.data, 3, A.KRUSKAL_L1, 0, None) < 0
    assert L.alto_kruskal_normalize(3, dims, 2, up, lam.ctypes.data, 0, 9, 0, None) < 0
    fit = ctypes.c_double()
    assert L.alto_cpd_fit_parts(3, dims, 2, None, lam.ctypes.data, up[2], up[2], 1.0, 0, ctypes.byref(fit)) < 0
    assert L.alto_cpd_fit_parts(0, dims, 2, up, lam.ctypes.data, up[2], up[2], 1.0, 0, ctypes.byref(fit)) < 0
    res = A.AprResult()
    assert L.alto_cp_apr_ex(None, 2, 1, 1, 0.0, 0.0, 0.0, 0.0, A.APR_STOP_COO, up, lam.ctypes.data,
                            ctypes.byref(res), 0, None) < 0
