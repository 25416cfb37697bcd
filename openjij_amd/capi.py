"""ctypes binding of the C ABI in include/ojmi.h (libojmi.so).

This is the same boundary a non-Python host would bind; the pybind11
``cxxjij`` module sits on it too.  Raises on any non-zero status.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libojmi.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "ojmi.h")

OK, EINVAL, ERANGE, EDEVICE, ENOMEM, ENOTSUP, ERUNTIME = range(7)
PREC_AUTO, PREC_FP64, PREC_FP32, PREC_INT32 = range(4)
KERNEL_DENSE_WAVE, KERNEL_SQA_WAVE, KERNEL_CSR_LEVEL, KERNEL_DENSE_SHARED = 1, 2, 3, 4
KERNEL_DENSE_COLOWN, KERNEL_BATCH_WAVE, KERNEL_DENSE_LDS, KERNEL_BATCH_PERSIST = 5, 6, 7, 8
KERNEL_BATCH_FUSED, KERNEL_DENSE_MFMA, KERNEL_SQA_CSR = 9, 10, 11


class Info(C.Structure):
    _fields_ = [
        ("num_spins", C.c_int64), ("num_replicas", C.c_int64), ("num_trotter", C.c_int64),
        ("precision", C.c_int32), ("kernel", C.c_int32), ("refresh_every", C.c_int32),
        ("coupling_bytes", C.c_int32), ("last_sweeps", C.c_int64),
        ("last_attempts", C.c_int64), ("last_kernel_ms", C.c_double),
        ("last_refresh_ms", C.c_double), ("group", C.c_int32), ("ring_rows", C.c_int32),
        ("lds_bytes", C.c_int32), ("lds_budget", C.c_int32), ("num_problems", C.c_int64),
        ("waves_per_replica", C.c_int32), ("mfma_segments", C.c_int32),
        ("last_refreshes", C.c_int64), ("last_flips", C.c_int64), ("last_rows", C.c_int64),
    ]


class OjmiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ojmi error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built")
        L = C.CDLL(LIB_PATH)
        vp, i64, u64, i32, dbl = C.c_void_p, C.c_int64, C.c_uint64, C.c_int32, C.c_double
        sig = {
            "ojmi_last_error": (C.c_char_p, []),
            "ojmi_device_count": (C.c_int, [C.POINTER(C.c_int)]),
            "ojmi_create_dense": (C.c_int, [i64, vp, i64, i64, dbl, i32, C.POINTER(vp)]),
            "ojmi_create_csr": (C.c_int, [i64, vp, vp, vp, i64, i64, dbl, i32, C.POINTER(vp)]),
            "ojmi_set_option": (C.c_int, [vp, C.c_char_p, i64]),
            "ojmi_set_spins": (C.c_int, [vp, vp]),
            "ojmi_get_spins": (C.c_int, [vp, vp]),
            "ojmi_set_gamma": (C.c_int, [vp, dbl]),
            "ojmi_run": (C.c_int, [vp, vp, vp, vp, i64, u64, u64]),
            "ojmi_solution": (C.c_int, [vp, vp]),
            "ojmi_energies": (C.c_int, [vp, vp]),
            "ojmi_field_drift": (C.c_int, [vp, C.POINTER(dbl)]),
            "ojmi_get_info": (C.c_int, [vp, C.POINTER(Info)]),
            "ojmi_destroy": (C.c_int, [vp]),
            "ojmi_create_batch": (C.c_int, [i64, vp, vp, i64, i32, C.POINTER(vp)]),
            "ojmi_run_batch": (C.c_int, [vp, vp, vp, i64, u64, u64]),
            "ojmi_refresh_fields": (C.c_int, [vp]),
            "ojmi_energies_static": (C.c_int, [vp, vp]),
            "ojmi_batch_beta_stats": (C.c_int, [vp, vp, vp]),
            "ojmi_get_fields": (C.c_int, [vp, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def header_symbols():
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(ojmi_[a-z0-9_]+)\s*\(", text)))


def _check(rc):
    if rc != OK:
        raise OjmiError(rc, lib().ojmi_last_error().decode())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    c = C.c_int(0)
    _check(lib().ojmi_device_count(C.byref(c)))
    return c.value


class Context:
    """One ojmi_ctx: R replicas (x M Trotter slices) of one problem."""

    def __init__(self, handle, n, R, M):
        self.h = handle
        self.n, self.R, self.M = n, R, M

    @classmethod
    def dense(cls, J_ext, R, M=1, gamma=0.0, precision=PREC_AUTO):
        J = np.ascontiguousarray(J_ext, dtype=np.float64)
        n = J.shape[0] - 1
        h = C.c_void_p()
        _check(lib().ojmi_create_dense(n, _p(J), R, M, gamma, precision, C.byref(h)))
        return cls(h, n, R, M)

    @classmethod
    def csr(cls, n, rowptr, col, val, R, M=1, gamma=0.0, precision=PREC_AUTO):
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        ci = np.ascontiguousarray(col, dtype=np.int64)
        v = np.ascontiguousarray(val, dtype=np.float64)
        h = C.c_void_p()
        _check(lib().ojmi_create_csr(n, _p(rp), _p(ci), _p(v), R, M, gamma, precision,
                                     C.byref(h)))
        return cls(h, n, R, M)

    def set_option(self, key, value):
        _check(lib().ojmi_set_option(self.h, key.encode(), int(value)))

    def set_spins(self, spins):
        s = np.ascontiguousarray(spins, dtype=np.int8)
        if s.size != self.R * self.M * self.n:
            raise ValueError("spins must hold R*M*n entries")
        _check(lib().ojmi_set_spins(self.h, _p(s)))

    def get_spins(self):
        s = np.empty((self.R, self.M, self.n) if self.M > 1 else (self.R, self.n), np.int8)
        _check(lib().ojmi_get_spins(self.h, _p(s)))
        return s

    def set_gamma(self, g):
        _check(lib().ojmi_set_gamma(self.h, float(g)))

    def run(self, beta, mcs, seed, sweep_offset=0, s=None):
        b = np.ascontiguousarray(beta, dtype=np.float64)
        m = np.ascontiguousarray(mcs, dtype=np.int64)
        sv = None if s is None else np.ascontiguousarray(s, dtype=np.float64)
        _check(lib().ojmi_run(self.h, _p(b), None if sv is None else _p(sv), _p(m), len(b),
                              int(seed), int(sweep_offset)))

    def solution(self):
        o = np.empty((self.R, self.n), np.int8)
        _check(lib().ojmi_solution(self.h, _p(o)))
        return o

    def energies(self):
        o = np.empty((self.R, self.M) if self.M > 1 else self.R)
        _check(lib().ojmi_energies(self.h, _p(o)))
        return o

    def drift(self):
        d = C.c_double(0)
        _check(lib().ojmi_field_drift(self.h, C.byref(d)))
        return d.value

    def refresh_fields(self):
        """Fields of every replica from its current spins on the matrix cores."""
        _check(lib().ojmi_refresh_fields(self.h))

    def get_fields(self):
        o = np.empty((self.R, self.n), np.float64)
        _check(lib().ojmi_get_fields(self.h, _p(o)))
        return o

    def info(self) -> dict:
        i = Info()
        _check(lib().ojmi_get_info(self.h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in Info._fields_}

    def close(self):
        if self.h:
            lib().ojmi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class BatchContext(Context):
    """Batch of K independent dense problems of sizes n[k], R replicas each
    (ojmi_create_batch).  Spins are a list of [R][n_k] arrays."""

    def __init__(self, handle, ns, R):
        super().__init__(handle, int(sum(ns)), R, 1)
        self.ns = [int(x) for x in ns]

    @classmethod
    def create(cls, J_exts, R, precision=PREC_AUTO):
        ns = np.array([np.asarray(J).shape[0] - 1 for J in J_exts], dtype=np.int64)
        flat = np.concatenate([np.ascontiguousarray(J, dtype=np.float64).ravel()
                               for J in J_exts])
        h = C.c_void_p()
        _check(lib().ojmi_create_batch(len(ns), _p(ns), _p(flat), R, precision, C.byref(h)))
        return cls(h, ns, R)

    def set_spins(self, spins):
        """spins: list of [R][n_k] +-1 arrays, one per problem."""
        flat = np.concatenate([np.ascontiguousarray(s, dtype=np.int8).ravel() for s in spins])
        super().set_spins(flat)

    def _split(self, flat):
        out, o = [], 0
        for n in self.ns:
            out.append(flat[o:o + self.R * n].reshape(self.R, n))
            o += self.R * n
        return out

    def get_spins(self):
        flat = np.empty(self.R * self.n, np.int8)
        _check(lib().ojmi_get_spins(self.h, _p(flat)))
        return self._split(flat)

    def solution(self):
        return self.get_spins()

    def run_batch(self, betas, mcs, seed, sweep_offset=0):
        """betas: [K][L] per-problem temperatures; mcs: [L] sweeps per entry."""
        b = np.ascontiguousarray(betas, dtype=np.float64)
        m = np.ascontiguousarray(mcs, dtype=np.int64)
        if b.ndim != 2 or b.shape != (len(self.ns), len(m)):
            raise ValueError("betas must be [num_problems][len(mcs)]")
        _check(lib().ojmi_run_batch(self.h, _p(b), _p(m), len(m), int(seed), int(sweep_offset)))

    def energies(self):
        e = np.empty((len(self.ns), self.R), np.float64)
        _check(lib().ojmi_energies(self.h, _p(e)))
        return e

    def beta_stats(self, spin_pairs):
        """spin_pairs: list of [2][n_k] +-1 arrays.  Returns [K][5]: dE_max,
        dE_min, lin_max, mean_abs, int_min of each problem (ojmi_batch_beta_stats)."""
        flat = np.concatenate([np.ascontiguousarray(s, dtype=np.int8).ravel()
                               for s in spin_pairs])
        if flat.size != 2 * sum(self.ns):
            raise ValueError("need two spin vectors of n_k entries per problem")
        out = np.empty((len(self.ns), 5), np.float64)
        _check(lib().ojmi_batch_beta_stats(self.h, _p(flat), _p(out)))
        return out
