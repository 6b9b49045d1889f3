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
