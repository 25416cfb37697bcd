"""ctypes front end of the CPU oracle (oracle/oj_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboj_oracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(LIB_PATH)
        i64, u64, u32, dbl, flt = C.c_int64, C.c_uint64, C.c_uint32, C.c_double, C.c_float
        vp = C.c_void_p
        sig = {
            "orc_philox4x32_10": (None, [vp, vp, vp]),
            "orc_philox_word": (u32, [u64, u64, u64, u64, u64]),
            "orc_det_log_d": (dbl, [dbl]),
            "orc_det_log_f": (flt, [flt]),
            "orc_theta_d": (dbl, [u32, dbl]),
            "orc_theta_f": (flt, [u32, flt]),
            "orc_theta_hb_d": (dbl, [u32, dbl]),
            "orc_theta_hb_f": (flt, [u32, flt]),
            "orc_ref_xorshift": (None, [C.c_uint, i64, vp]),
            "orc_ref_uniform": (None, [C.c_uint, i64, vp]),
            "orc_make_classical_schedule": (None, [dbl, dbl, i64, vp]),
            "orc_make_transverse_schedule": (None, [i64, vp]),
            "orc_ref_gen_spin": (None, [i64, u64, vp]),
            "orc_expand_transverse": (i64, [vp, vp, vp, i64, dbl, i64, vp, vp]),
            "orc_ref_sa_dense": (C.c_int, [i64, vp, vp, u64, vp, vp, i64]),
            "orc_ref_sa_dense_batch": (C.c_int, [i64, vp, i64, vp, vp, vp, vp, i64, C.c_int]),
            "orc_ref_sa_csr": (C.c_int, [i64, vp, vp, vp, vp, u64, vp, vp, i64]),
            "orc_ref_sa_csr_batch": (C.c_int, [i64, vp, vp, vp, i64, vp, vp, vp, vp, i64, C.c_int]),
            "orc_ref_sqa_dense": (C.c_int, [i64, vp, i64, dbl, vp, u64, vp, vp, vp, i64]),
            "orc_ref_solution_transverse": (C.c_int, [i64, vp, i64, vp, vp]),
            "orc_energy_exact_dense": (dbl, [i64, vp, vp]),
            "orc_energy_exact_dense_batch": (None, [i64, vp, i64, vp, vp]),
            "orc_energy_exact_csr": (dbl, [i64, vp, vp, vp, vp]),
            "orc_brute_force_dense": (C.c_int, [i64, vp, vp, vp]),
            "orc_replay_dense": (C.c_int, [C.c_int, i64, vp, i64, vp, u64, u64, vp, i64]),
            "orc_replay_dense_w": (C.c_int, [C.c_int, i64, vp, i64, vp, u64, u64, vp, i64, u32]),
            "orc_replay_dense_wr": (C.c_int, [C.c_int, i64, vp, i64, vp, u64, u64, vp, i64, u32,
                                              C.c_int]),
            "orc_replay_dense_from_fields": (C.c_int, [C.c_int, i64, vp, i64, vp, vp, u64, u64,
                                                       vp, i64, u32, C.c_int]),
            "orc_replay_csr": (C.c_int, [C.c_int, i64, vp, vp, vp, i64, vp, u64, u64, vp, i64]),
            "orc_replay_csr_r": (C.c_int, [C.c_int, i64, vp, vp, vp, i64, vp, u64, u64, vp, i64,
                                           C.c_int]),
            "orc_replay_sqa_dense": (C.c_int, [C.c_int, i64, vp, i64, i64, vp, u64, u64, vp, vp, i64]),
            "orc_replay_sqa_csr": (C.c_int, [C.c_int, i64, vp, vp, vp, i64, i64, vp, u64, u64, vp, vp, i64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _i8(a):
    return np.ascontiguousarray(a, dtype=np.int8)


MODE = {"fp64": 0, "fp32": 1, "int32": 2}


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def philox_word(seed, t, r, m, i):
    return lib().orc_philox_word(seed, t, r, m, i)


def det_log_d(x):
    return lib().orc_det_log_d(float(x))


def det_log_f(x):
    return lib().orc_det_log_f(float(np.float32(x)))


def ref_xorshift(seed, count):
    o = np.zeros(count, dtype=np.uint32)
    lib().orc_ref_xorshift(seed, count, _p(o))
    return o


def ref_uniform(seed, count):
    o = np.zeros(count)
    lib().orc_ref_uniform(seed, count, _p(o))
    return o


def classical_schedule(beta_min, beta_max, L):
    o = np.zeros(L)
    lib().orc_make_classical_schedule(beta_min, beta_max, L, _p(o))
    return o


def transverse_schedule(L):
    o = np.zeros(L)
    lib().orc_make_transverse_schedule(L, _p(o))
    return o


def ref_gen_spin(n, seed):
    o = np.zeros(n, dtype=np.int8)
    lib().orc_ref_gen_spin(n, seed, _p(o))
    return o


def ref_sa_dense(J_ext, spins, seed, beta, mcs):
    J = _f64(J_ext)
    s = _i8(spins).copy()
    b, m = _f64(beta), _i64(mcs)
    rc = lib().orc_ref_sa_dense(J.shape[0] - 1, _p(J), _p(s), seed, _p(b), _p(m), len(b))
    assert rc == 0
    return s


def ref_sa_dense_batch(J_ext, spins, seeds, beta, mcs, nthreads=0):
    J = _f64(J_ext)
    s = _i8(spins).copy()
    sd = np.ascontiguousarray(seeds, dtype=np.uint32)
    b, m = _f64(beta), _i64(mcs)
    rc = lib().orc_ref_sa_dense_batch(J.shape[0] - 1, _p(J), s.shape[0], _p(s), _p(sd), _p(b),
                                      _p(m), len(b), nthreads)
    assert rc == 0
    return s


def ref_sa_csr_batch(rowptr, col, val, n, spins, seeds, beta, mcs, nthreads=0):
    rp, ci, v = _i64(rowptr), _i64(col), _f64(val)
    s = _i8(spins).copy()
    sd = np.ascontiguousarray(seeds, dtype=np.uint32)
    b, m = _f64(beta), _i64(mcs)
    rc = lib().orc_ref_sa_csr_batch(n, _p(rp), _p(ci), _p(v), s.shape[0], _p(s), _p(sd), _p(b),
                                    _p(m), len(b), nthreads)
    assert rc == 0
    return s


def ref_sqa_dense(J_ext, gamma, tspins, seed, beta, s, mcs):
    J = _f64(J_ext)
    t = _i8(tspins).copy()
    b, sv, m = _f64(beta), _f64(s), _i64(mcs)
    M = t.shape[0]
    rc = lib().orc_ref_sqa_dense(J.shape[0] - 1, _p(J), M, gamma, _p(t), seed, _p(b), _p(sv),
                                 _p(m), len(b))
    assert rc == 0
    return t


def ref_solution_transverse(J_ext, tspins):
    J = _f64(J_ext)
    t = _i8(tspins)
    o = np.zeros(t.shape[1], dtype=np.int8)
    lib().orc_ref_solution_transverse(J.shape[0] - 1, _p(J), t.shape[0], _p(t), _p(o))
    return o


def energy_exact(J_ext, spins):
    J = _f64(J_ext)
    s = _i8(spins)
    n = J.shape[0] - 1
    if s.ndim == 1:
        return lib().orc_energy_exact_dense(n, _p(J), _p(s))
    out = np.zeros(s.shape[0])
    lib().orc_energy_exact_dense_batch(n, _p(J), s.shape[0], _p(s), _p(out))
    return out


def energy_exact_csr(rowptr, col, val, n, spins):
    rp, ci, v = _i64(rowptr), _i64(col), _f64(val)
    s = _i8(spins)
    return lib().orc_energy_exact_csr(n, _p(rp), _p(ci), _p(v), _p(s))


def brute_force(J_ext):
    J = _f64(J_ext)
    n = J.shape[0] - 1
    e = C.c_double(0)
    st = np.zeros(n, dtype=np.int8)
    rc = lib().orc_brute_force_dense(n, _p(J), C.byref(e), _p(st))
    assert rc == 0
    return e.value, st


def replay_dense(mode, J_ext, spins, seed, sweep_offset, nbinv, word3=0, rule=0):
    """word3: fourth Philox counter word (the problem index inside a batch);
    rule: 0 = Metropolis, 1 = heat bath."""
    J = _f64(J_ext)
    s = _i8(spins).copy()
    nb = _f64(nbinv)
    rc = lib().orc_replay_dense_wr(MODE[mode], J.shape[0] - 1, _p(J), s.shape[0], _p(s), seed,
                                   sweep_offset, _p(nb), len(nb), word3, rule)
    assert rc == 0, rc
    return s


def replay_dense_from_fields(mode, J_ext, spins, fields, seed, sweep_offset, nbinv, word3=0,
                             rule=0):
    """replay_dense starting every replica from the given fields [R][n]
    (instead of accumulating them from the spins)."""
    J = _f64(J_ext)
    s = _i8(spins).copy()
    F = _f64(fields)
    assert F.shape == s.shape
    nb = _f64(nbinv)
    rc = lib().orc_replay_dense_from_fields(MODE[mode], J.shape[0] - 1, _p(J), s.shape[0], _p(s),
                                            _p(F), seed, sweep_offset, _p(nb), len(nb), word3,
                                            rule)
    assert rc == 0, rc
    return s


def replay_csr(mode, rowptr, col, val, n, spins, seed, sweep_offset, nbinv, rule=0):
    rp, ci, v = _i64(rowptr), _i64(col), _f64(val)
    s = _i8(spins).copy()
    nb = _f64(nbinv)
    rc = lib().orc_replay_csr_r(MODE[mode], n, _p(rp), _p(ci), _p(v), s.shape[0], _p(s), seed,
                                sweep_offset, _p(nb), len(nb), rule)
    assert rc == 0, rc
    return s


def replay_sqa_dense(mode, J_ext, spins, seed, sweep_offset, a, B):
    J = _f64(J_ext)
    s = _i8(spins).copy()  # R x M x n
    av, Bv = _f64(a), _f64(B)
    rc = lib().orc_replay_sqa_dense(MODE[mode], J.shape[0] - 1, _p(J), s.shape[0], s.shape[1],
                                    _p(s), seed, sweep_offset, _p(av), _p(Bv), len(av))
    assert rc == 0, rc
    return s


def replay_sqa_csr(mode, rowptr, col, val, n, spins, seed, sweep_offset, a, B):
    rp, ci, v = _i64(rowptr), _i64(col), _f64(val)
    s = _i8(spins).copy()
    av, Bv = _f64(a), _f64(B)
    rc = lib().orc_replay_sqa_csr(MODE[mode], n, _p(rp), _p(ci), _p(v), s.shape[0], s.shape[1],
                                  _p(s), seed, sweep_offset, _p(av), _p(Bv), len(av))
    assert rc == 0, rc
    return s


# --- helpers shared by the tests -------------------------------------------
def expand_classical(beta, mcs):
    """Per-sweep -1/beta as the engine computes it."""
    out = []
    for b, m in zip(beta, mcs):
        out += [-1.0 / b] * int(m)
    return np.array(out)


def expand_transverse(beta, s, mcs, gamma, M):
    """Per-sweep (a, B) of the Trotter update, computed with the C library's
    log/tanh as the engine's host code does."""
    b, sv, m = _f64(beta), _f64(s), _i64(mcs)
    T = int(m.sum())
    a = np.zeros(T)
    B = np.zeros(T)
    lib().orc_expand_transverse(_p(b), _p(sv), _p(m), len(b), float(gamma), int(M), _p(a), _p(B))
    return a, B


def dense_to_csr(J_ext):
    """Full symmetric (n+1)-row CSR with sorted columns (zeros dropped)."""
    J = np.asarray(J_ext, dtype=float)
    n1 = J.shape[0]
    S = np.triu(J) + np.triu(J, 1).T
    rowptr = [0]
    col, val = [], []
    for i in range(n1):
        nz = np.flatnonzero(S[i])
        col += list(nz)
        val += list(S[i, nz])
        rowptr.append(len(col))
    return np.array(rowptr), np.array(col), np.array(val)
