"""The dense sweep whose row updates run on the matrix cores (option
``mfma_sweep``, kernel k_dense_mfma): 8 replicas per workgroup, each 64-spin
block decided in sweep order, then the block's flips applied to every field
as one contraction dF = C J_block.

INT32 fields with int8 couplings: the contraction is exact (int8 x int8 into
int32), so trajectories equal the oracle replay bit for bit.  That pins
everything except the FP32 arithmetic: Philox counters, block order, the
decision chains, the accumulator layout of the fields, spin storage, ragged n
and R, split runs and the heat-bath rule.

FP32 fields (bf16 J_hi + J_lo, FP32 accumulation): each block adds
sum_i c_i (J_hi + J_lo)[i][j] with one FP32 rounding per matrix instruction
instead of one per flip, so the fields differ from the sequential FP32
replay in the last bits.  Checked: (a) the fields a run leaves against the
exact FP64 fields of its final spins (stated tolerance, printed with the
measured value); (b) trajectories against the FP32 replay from the same
initial fields, replica by replica (a replica can leave the replay's path only
where a flip test lands within a few ulp of its threshold); (c) the final
energy distribution of an anneal against the plain FP64 path (two-sample KS).
"""
import numpy as np
import pytest

from openjij_amd import capi
from oracle import oracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu

PREC = {"fp64": capi.PREC_FP64, "fp32": capi.PREC_FP32, "int32": capi.PREC_INT32}


def exact_fields(E, spins):
    S = H.symmetric_ext(E)
    n = S.shape[0] - 1
    return spins.astype(np.float64) @ S[:n, :n] + S[:n, n]


def mfma_ctx(E, R, prec, form=0):
    """form 0: every wave decides and updates; 1: four waves hold the fields
    and stream the updates while four decide."""
    ctx = capi.Context.dense(E, R, 1, 0.0, PREC[prec])
    ctx.set_option("mfma_sweep", 1)
    ctx.set_option("mfma_form", form)
    info = ctx.info()
    assert info["kernel"] == capi.KERNEL_DENSE_MFMA and info["group"] == 8
    return ctx


@pytest.mark.parametrize("n,R,L", [(1000, 8, 10), (1025, 10, 8), (2048, 17, 6), (4096, 16, 4),
                                   (3000, 3, 5), (961, 9, 12)])
@pytest.mark.parametrize("rule", [0, 1])
@pytest.mark.parametrize("form", [0, 1])
def test_int32_bitexact(n, R, L, rule, form):
    print(f"mfma sweep int32: n={n} R={R} sweeps={L} rule={rule} form={form}", flush=True)
    E = H.int_dense(n, 0x3A00 + n)
    beta, mcs = H.geometric_betas(0.05, 3.0, L), np.ones(L, dtype=np.int64)
    init = H.random_pm1(R, n, n + 3)
    with mfma_ctx(E, R, "int32", form) as ctx:
        ctx.set_option("update_method", rule)
        ctx.set_spins(init)
        ctx.run(beta, mcs, 0x5EED, 7)
        got = ctx.get_spins()
        info = ctx.info()
        en = ctx.energies()
    want = orc.replay_dense("int32", H.symmetric_ext(E), init, 0x5EED, 7,
                            orc.expand_classical(beta, mcs), rule=rule)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(en, orc.energy_exact(E, got))
    # every flip counted once; every block of every workgroup reads its 64 rows
    assert info["last_flips"] > 0
    assert info["last_rows"] == ((R + 7) // 8) * L * n


@pytest.mark.parametrize("form", [0, 1])
def test_int32_split_run_and_pacing(form):
    """Two runs with consecutive sweep offsets equal one run; 32 workgroups,
    so XCD pacing is active (performance only)."""
    n, R = 1024, 256
    E = H.int_dense(n, 0x3B00)
    beta, mcs = H.geometric_betas(0.1, 4.0, 12), np.ones(12, dtype=np.int64)
    init = H.random_pm1(R, n, 11)
    with mfma_ctx(E, R, "int32", form) as ctx:
        ctx.set_spins(init)
        ctx.run(beta[:5], mcs[:5], 99, 0)
        ctx.run(beta[5:], mcs[5:], 99, 5)
        got = ctx.get_spins()
    want = orc.replay_dense("int32", H.symmetric_ext(E), init, 99, 0,
                            orc.expand_classical(beta, mcs))
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("n,R", [(4096, 16), (1024, 24), (2000, 8)])
@pytest.mark.parametrize("form", [0, 1])
def test_fp32_fields_and_trajectories(n, R, form):
    """FP32: fields after a run against the exact fields of the final spins,
    and trajectories against the FP32 replay from the same initial fields."""
    print(f"mfma sweep fp32: n={n} R={R}", flush=True)
    E = H.sk_gauss(n, 0x3C00 + n, field=True)
    L = 4
    beta, mcs = H.geometric_betas(0.5, 2.0, L), np.ones(L, dtype=np.int64)
    init = H.random_pm1(R, n, n + 5)
    with capi.Context.dense(E, R, 1, 0.0, capi.PREC_FP32) as ref:
        ref.set_spins(init)
        ref.refresh_fields()
        F0 = ref.get_fields()
    with mfma_ctx(E, R, "fp32", form) as ctx:
        ctx.set_spins(init)
        ctx.run(beta, mcs, 0xF00D, 0)
        got = ctx.get_spins()
        F = ctx.get_fields()
        drift = ctx.drift()
    err = np.abs(F - exact_fields(E, got))
    scale = np.abs(H.symmetric_ext(E)[:n]).sum(axis=1)  # |h_i| + sum_j |J_ij|
    print(f"  max |F - F_exact| = {err.max():.3e}, max field scale {scale.max():.2f}, "
          f"GPU drift {np.max(drift):.3e}")
    # stated tolerance: 2e-6 of the field scale after 4 sweeps (measured
    # values are printed above; FP32 with per-block rounding stays near
    # 1e-7 of it)
    assert np.all(err <= 2e-6 * scale[None, :])
    want = orc.replay_dense_from_fields("fp32", H.symmetric_ext(E), init, F0, 0xF00D, 0,
                                        orc.expand_classical(beta, mcs))
    same = np.all(got == want, axis=1)
    print(f"  replicas identical to the sequential FP32 replay: {same.sum()}/{R}")
    assert same.mean() >= 0.75


def test_fp32_anneal_energy_distribution_matches_fp64():
    """Final energies of an anneal: matrix-core FP32 path against the plain
    FP64 path on the same instance and schedule (two-sample KS)."""
    n, R, L = 1024, 512, 60
    E = H.sk_gauss(n, 0x3D00)
    beta, mcs = H.geometric_betas(0.1, 8.0, L), np.ones(L, dtype=np.int64)
    init = H.random_pm1(R, n, 21)
    with mfma_ctx(E, R, "fp32") as ctx:
        ctx.set_option("refresh_every", 16)
        ctx.set_spins(init)
        ctx.run(beta, mcs, 1, 0)
        e32 = ctx.energies()
        s32 = ctx.get_spins()
    with capi.Context.dense(E, R, 1, 0.0, capi.PREC_FP64) as ctx:
        ctx.set_spins(init)
        ctx.run(beta, mcs, 2, 0)
        e64 = ctx.energies()
    np.testing.assert_allclose(e32, orc.energy_exact(E, s32), rtol=1e-12, atol=1e-9)
    p = H.ks_2samp_pvalue(e32, e64)
    print(f"KS p = {p:.3f}, mean {e32.mean():.4f} vs {e64.mean():.4f}")
    assert p > 1e-3


def test_option_rejected_where_unsupported():
    E = H.sk_gauss(500, 5)  # 8 slots of 64 spins: below the kernel's 16
    with capi.Context.dense(E, 8, 1, 0.0, capi.PREC_FP32) as ctx:
        with pytest.raises(capi.OjmiError):
            ctx.set_option("mfma_sweep", 1)
    with capi.Context.dense(H.sk_gauss(1000, 6), 8, 1, 0.0, capi.PREC_FP64) as ctx:
        with pytest.raises(capi.OjmiError):
            ctx.set_option("mfma_sweep", 1)


def test_int32_hybrid_segments_bitexact():
    """mfma_sweep = 2: each 16-sweep segment runs the matrix-core sweep or the
    column-owned kernel by the acceptance rate the GPU counted in the segment
    before.  Both are exact for INT32 fields, so the trajectory stays the
    oracle replay's whichever ran."""
    n, R = 1024, 64
    E = H.int_dense(n, 0x3E00)
    L = 64
    beta, mcs = H.geometric_betas(0.05, 4.0, L), np.ones(L, dtype=np.int64)
    init = H.random_pm1(R, n, 31)
    with capi.Context.dense(E, R, 1, 0.0, capi.PREC_INT32) as ctx:
        ctx.set_option("group", 4)
        ctx.set_option("mfma_sweep", 2)
        ctx.set_option("mfma_accept", 300)  # switches within this schedule
        assert ctx.info()["kernel"] == capi.KERNEL_DENSE_MFMA
        ctx.set_spins(init)
        ctx.run(beta, mcs, 0xAB, 3)
        got = ctx.get_spins()
        segs = ctx.info()["mfma_segments"]
    print(f"segments on the matrix-core sweep: {segs} of 4")
    assert 0 < segs < 4  # both kernels ran
    want = orc.replay_dense("int32", H.symmetric_ext(E), init, 0xAB, 3,
                            orc.expand_classical(beta, mcs))
    np.testing.assert_array_equal(got, want)


def test_fp32_default_large_context_is_hybrid_and_replays():
    """Dense FP32 with 2048 replicas and 4096 spins (the SK N = 4096 shape):
    the default is mfma_sweep = 2 with a 16-sweep field refresh.  A hot
    segment (matrix-core sweep) then two cold ones (column-owned kernel)
    against the FP32 replay of 16 of the replicas, each segment replayed from
    the fields the GPU's refresh gives for its starting spins."""
    from tests.test_gpu_mfma import chained_replay

    n, R, K = 4096, 2048, 16
    E = H.sk_gauss(n, 0xE005)
    beta = np.concatenate([np.full(K, 0.2), np.full(2 * K, 6.0)])
    mcs = np.ones(len(beta), dtype=np.int64)
    init = H.random_pm1(R, n, 41)
    with capi.Context.dense(E, R, 1, 0.0, capi.PREC_FP32) as ctx:
        info = ctx.info()
        assert info["kernel"] == capi.KERNEL_DENSE_MFMA and info["group"] == 8
        ctx.set_option("refresh_every", K)
        ctx.set_spins(init)
        ctx.run(beta, mcs, 0x5151, 0)
        got = ctx.get_spins()[:16]
        F = ctx.get_fields()[:16]
        en = ctx.energies()
        segs = ctx.info()["mfma_segments"]
    print(f"segments on the matrix-core sweep: {segs} of 3")
    assert 0 < segs < 3  # hot segments on the matrix cores, the last cold one not
    with capi.Context.dense(E, 16, 1, 0.0, capi.PREC_FP32) as refresher:
        want = chained_replay(E, init[:16], beta, mcs, K, 0x5151, 0, refresher)
    same = np.all(got == want, axis=1)
    print(f"replicas identical to the segment-wise FP32 replay: {same.sum()}/16")
    assert same.mean() >= 0.75
    err = np.abs(F - exact_fields(E, got))
    scale = np.abs(H.symmetric_ext(E)[:n]).sum(axis=1)
    print(f"max |F - F_exact| = {err.max():.3e}")
    assert np.all(err <= 2e-6 * scale[None, :])
    np.testing.assert_allclose(en[:16], orc.energy_exact(E, got),
                               rtol=1e-12, atol=1e-9)
