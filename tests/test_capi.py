"""The C-ABI library loads, exports every symbol of include/ojmi.h and
validates its arguments before touching a device.  CPU only: no compute call."""
import ctypes as C

import numpy as np
import pytest

from openjij_amd import capi


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = capi.header_symbols()
    assert len(names) >= 14
    for name in names:
        assert hasattr(L, name), name


def test_device_count_without_gpu_is_callable():
    assert capi.device_count() >= 0


@pytest.mark.parametrize("n,R,M,code", [(0, 1, 1, capi.EINVAL), (4, 0, 1, capi.EINVAL),
                                         (4, 1, 0, capi.EINVAL)])
def test_create_dense_rejects_bad_shapes(n, R, M, code):
    J = np.zeros((max(n, 1) + 1, max(n, 1) + 1))
    h = C.c_void_p()
    rc = capi.lib().ojmi_create_dense(n, J.ctypes.data_as(C.c_void_p), R, M, 0.0, 0,
                                      C.byref(h))
    assert rc == code
    assert capi.lib().ojmi_last_error()


def test_create_dense_rejects_nonzero_diagonal():
    J = np.zeros((4, 4))
    J[1, 1] = 2.0
    J[3, 3] = 1.0
    h = C.c_void_p()
    rc = capi.lib().ojmi_create_dense(3, J.ctypes.data_as(C.c_void_p), 1, 1, 0.0, 0, C.byref(h))
    assert rc == capi.ERUNTIME
    assert b"diagonal" in capi.lib().ojmi_last_error()


def test_create_dense_rejects_int_precision_for_real_couplings():
    J = np.zeros((4, 4))
    J[0, 1] = 0.5
    J[3, 3] = 1.0
    h = C.c_void_p()
    rc = capi.lib().ojmi_create_dense(3, J.ctypes.data_as(C.c_void_p), 1, 1, 0.0,
                                      capi.PREC_INT32, C.byref(h))
    assert rc == capi.EINVAL


def test_null_context_calls_fail_cleanly():
    L = capi.lib()
    assert L.ojmi_set_spins(None, None) == capi.EINVAL
    assert L.ojmi_run(None, None, None, None, 0, 0, 0) == capi.EINVAL
    assert L.ojmi_destroy(None) == capi.OK


def _batch(ns, R=1, prec=0):
    """ojmi_create_batch on empty problems of sizes ns (corner 1 only)."""
    blocks = []
    for n in ns:
        b = np.zeros((max(n, 0) + 1, max(n, 0) + 1))
        b[-1, -1] = 1.0
        blocks.append(b.ravel())
    flat = np.concatenate(blocks) if blocks else np.zeros(1)
    nn = np.array(ns if ns else [0], dtype=np.int64)
    h = C.c_void_p()
    rc = capi.lib().ojmi_create_batch(len(ns), nn.ctypes.data_as(C.c_void_p),
                                      flat.ctypes.data_as(C.c_void_p), R, prec, C.byref(h))
    return rc


@pytest.mark.parametrize("ns,R,code", [([3, 0, 2], 1, capi.EINVAL), ([3, 2], 0, capi.EINVAL),
                                        ([5000], 1, capi.ENOTSUP), ([], 1, capi.EINVAL)])
def test_create_batch_rejects_bad_shapes(ns, R, code):
    assert _batch(ns, R) == code
    assert capi.lib().ojmi_last_error()


def test_create_batch_rejects_bad_problem_and_names_it():
    blocks = [np.zeros((3, 3)), np.zeros((4, 4))]
    blocks[0][2, 2] = blocks[1][3, 3] = 1.0
    blocks[1][1, 1] = 0.5  # non-zero diagonal in problem 1
    flat = np.concatenate([b.ravel() for b in blocks])
    nn = np.array([2, 3], dtype=np.int64)
    h = C.c_void_p()
    rc = capi.lib().ojmi_create_batch(2, nn.ctypes.data_as(C.c_void_p),
                                      flat.ctypes.data_as(C.c_void_p), 1, 0, C.byref(h))
    assert rc == capi.ERUNTIME
    assert b"problem 1" in capi.lib().ojmi_last_error()


def test_run_batch_needs_a_batch_context():
    assert capi.lib().ojmi_run_batch(None, None, None, 0, 0, 0) == capi.EINVAL
