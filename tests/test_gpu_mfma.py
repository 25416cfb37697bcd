"""Matrix-core contractions on the GPU: local fields H = S J (bf16 hi/lo with
FP32 accumulation for FP32 contexts, exact int8 for INT32 ones), the periodic
field refresh of FP32 runs, FP64 energies, and the run counters.

Refresh parity: a run with refresh_every = K computes fields on the matrix
cores at sweeps 0, K, 2K, ... of the run and sweeps from them.  The oracle
replays each K-sweep segment bit for bit from the same fields, downloaded
from a second context that refreshes the same spins (the field kernel is
deterministic), so the whole run is pinned segment by segment.

Refresh accuracy bound (stated): with J_hi = bf16(J) and J_lo = bf16(J - J_hi),
|J - J_hi - J_lo| <= 2^-18 |J|; FP32 accumulation of the 2 n products
(exact in FP32: +-1 times bf16) in MFMA blocks of 32 adds at most one FP32
rounding per block and per half, so
  |F_gpu - F_exact| <= (2^-18 + (2 ceil(n/32) + 2) 2^-24) * (|h| + sum_j |J_ij|).
"""
import numpy as np
import pytest

from openjij_amd import capi
from oracle import oracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu


def exact_fields(E, spins):
    S = H.symmetric_ext(E)
    n = S.shape[0] - 1
    return spins.astype(np.float64) @ S[:n, :n] + S[:n, n]


def refresh_bound(E):
    S = H.symmetric_ext(E)
    n = S.shape[0] - 1
    row = np.abs(S[:n, :n]).sum(axis=1) + np.abs(S[:n, n])
    return (2.0 ** -18 + (2 * ((n + 31) // 32) + 2) * 2.0 ** -24) * row


@pytest.mark.parametrize("n,R", [(5, 3), (64, 8), (300, 257), (1024, 40), (4096, 16), (4500, 3)])
def test_fp32_refresh_fields_within_bound(n, R):
    E = H.sk_gauss(n, 0xF1 + n, field=True)
    init = H.random_pm1(R, n, n)
    with capi.Context.dense(E, R, 1, 0.0, capi.PREC_FP32) as ctx:
        ctx.set_spins(init)
        ctx.refresh_fields()
        F = ctx.get_fields()
    ex = exact_fields(E, init)
    err = np.abs(F - ex)
    bound = refresh_bound(E)
    print(f"n={n} R={R}: max |F - F_exact| = {err.max():.3e} (bound {bound.max():.3e})")
    assert np.all(err <= bound[None, :])


@pytest.mark.parametrize("n,R", [(7, 2), (200, 70), (700, 9), (4096, 12)])
def test_int8_fields_exact(n, R):
    E = H.int_dense(n, 0xF2 + n)
    init = H.random_pm1(R, n, n + 1)
    with capi.Context.dense(E, R, 1, 0.0, capi.PREC_INT32) as ctx:
        assert ctx.info()["coupling_bytes"] == 1
        ctx.set_spins(init)
        ctx.refresh_fields()
        F = ctx.get_fields()
    np.testing.assert_array_equal(F, exact_fields(E, init))


def chained_replay(E, init, beta, mcs, K, seed, offset, refresher):
    """Segments of K sweeps, each replayed from the fields the GPU's field
    kernel gives for the segment's starting spins."""
    nb = orc.expand_classical(beta, mcs)
    spins = init.copy()
    S = H.symmetric_ext(E)
    for t in range(0, len(nb), K):
        refresher.set_spins(spins)
        refresher.refresh_fields()
        F = refresher.get_fields()
        spins = orc.replay_dense_from_fields("fp32", S, spins, F, seed, offset + t, nb[t:t + K])
    return spins


@pytest.mark.parametrize("n,R,K,L", [(4096, 8, 2, 5), (700, 33, 3, 9), (100, 5, 4, 12),
                                     (4500, 2, 2, 3)])
def test_fp32_refresh_run_bitexact_per_segment(n, R, K, L):
    E = H.sk_gauss(n, 0xF3 + n)
    beta = H.geometric_betas(0.1, 10.0, 500)[:L] if n >= 4096 else H.geometric_betas(0.1, 6.0, L)
    mcs = np.ones(L, np.int64)
    init = H.random_pm1(R, n, 0xE1)
    with capi.Context.dense(E, R, 1, 0.0, capi.PREC_FP32) as ctx, \
            capi.Context.dense(E, R, 1, 0.0, capi.PREC_FP32) as ref:
        ctx.set_option("refresh_every", K)
        ctx.set_spins(init)
        ctx.run(beta, mcs, 4242, 7)
        got = ctx.get_spins()
        info = ctx.info()
        want = chained_replay(E, init, beta, mcs, K, 4242, 7, ref)
    assert info["refresh_every"] == K
    assert info["last_refreshes"] == (L + K - 1) // K
    assert 0 < info["last_refresh_ms"] < info["last_kernel_ms"]
    assert not np.array_equal(got, init)
    np.testing.assert_array_equal(got, want)


def test_refresh_reduces_field_drift():
    """Config E's shape, 40 sweeps: drift of the carried FP32 fields against
    FP64 fields of the final spins, without refresh and with refresh every 16
    sweeps (the last segment is 8 sweeps long)."""
    n, R, L = 4096, 64, 40
    E = H.sk_gauss(n, 0xE005)
    beta = H.geometric_betas(0.1, 10.0, 500)[:L]
    mcs = np.ones(L, np.int64)
    init = H.random_pm1(R, n, 3)
    drifts = {}
    for K in (0, 16):
        with capi.Context.dense(E, R, 1, 0.0, capi.PREC_FP32) as ctx:
            ctx.set_option("refresh_every", K)
            ctx.set_spins(init)
            ctx.run(beta, mcs, 9, 0)
            drifts[K] = ctx.drift()
            en = ctx.energies()
            F = ctx.get_fields()
            got = ctx.get_spins()
        ex = exact_fields(E, got)
        assert abs(np.abs(F - ex).max() - drifts[K]) <= 1e-12
        np.testing.assert_allclose(en, orc.energy_exact(E, got), rtol=1e-12, atol=1e-9)
    print("drift without / with refresh:", drifts)
    assert drifts[16] < drifts[0]


@pytest.mark.parametrize("prec,n,R,M", [("fp64", 77, 1, 1), ("fp32", 300, 257, 1),
                                        ("fp64", 130, 3, 5), ("int32", 1000, 65, 1),
                                        ("fp32", 4096, 8, 1)])
def test_energies_on_matrix_cores(prec, n, R, M):
    E = H.int_dense(n, 0xF4 + n) if prec == "int32" else H.sk_gauss(n, 0xF4 + n, field=True)
    P = {"fp64": capi.PREC_FP64, "fp32": capi.PREC_FP32, "int32": capi.PREC_INT32}[prec]
    init = H.random_pm1(R * M, n, 2).reshape((R, M, n) if M > 1 else (R, n))
    with capi.Context.dense(E, R, M, 1.0, P) as ctx:
        ctx.set_spins(init)
        en = ctx.energies()
    ex = orc.energy_exact(E, init.reshape(R * M, n)).reshape(en.shape)
    if prec == "int32":
        np.testing.assert_array_equal(en, ex)
    else:
        assert np.all(np.abs(en - ex) <= 1e-12 * (1 + np.abs(ex)))


@pytest.mark.parametrize("kind", ["wave", "colown", "shared", "lds", "csr", "sqa", "batch"])
def test_flip_counters_one_sweep(kind):
    """Each spin is attempted once per sweep, so after one sweep the flip
    counter equals the number of spins that changed; row-streaming kernels
    read at most one row per flip per replica."""
    beta, mcs = np.array([0.7]), np.ones(1, np.int64)
    if kind == "csr":
        n, rp, ci, v = H.ea3d_csr(6, 1)
        init = H.random_pm1(20, n, 4)
        ctx = capi.Context.csr(n, rp, ci, v, 20, precision=capi.PREC_INT32)
    elif kind == "batch":
        Es = [H.int_dense(m, 10 + m) for m in (40, 130, 300)]
        init = [H.random_pm1(4, E.shape[0] - 1, 5 + k) for k, E in enumerate(Es)]
        ctx = capi.BatchContext.create(Es, 4, capi.PREC_INT32)
    elif kind == "sqa":
        E = H.sk_gauss(100, 3)
        init = H.random_pm1(12, 100, 1).reshape(3, 4, 100)
        ctx = capi.Context.dense(E, 3, 4, 1.0, capi.PREC_FP32)
    else:
        n, R, prec = {"wave": (300, 5, capi.PREC_FP64), "colown": (600, 16, capi.PREC_FP32),
                      "shared": (300, 24, capi.PREC_FP32), "lds": (4200, 2, capi.PREC_INT32)}[kind]
        E = H.int_dense(n, n) if prec == capi.PREC_INT32 else H.sk_gauss(n, n)
        init = H.random_pm1(R, n, 6)
        ctx = capi.Context.dense(E, R, 1, 0.0, prec)
        if kind == "shared":
            ctx.set_option("group", 8)
            ctx.set_option("colown", 0)
    with ctx:
        ctx.set_spins(init)
        if kind == "sqa":
            ctx.run(np.array([2.0]), mcs, 1, 0, s=np.array([0.5]))
        else:
            ctx.run(beta, mcs, 1, 0)
        got = ctx.get_spins()
        info = ctx.info()
    changed = sum(int((np.asarray(g) != np.asarray(i)).sum()) for g, i in
                  (zip(got, init) if kind == "batch" else [(got, init)]))
    assert changed > 0
    assert info["last_flips"] == changed
    if kind in ("wave", "lds", "sqa", "batch"):
        assert info["last_rows"] == changed
    elif kind in ("colown", "shared"):
        assert 0 < info["last_rows"] <= changed
