"""GPU parity of the HIP kernels through the C ABI, against the CPU oracle.

Bit-exact: the Philox replay reproduces the GPU spin trajectories for integer
couplings on every path, and for real couplings on the fp64 and fp32 paths
(fields are accumulated in the same order on both sides).  Energies are
compared with the exact FP64 evaluation (equal for integer couplings,
relative 1e-12 for real ones).  Ground states are checked against the
reference fixtures and brute force; distributions against the reference
restatement with a two-sample KS test (threshold p > 1e-3).
"""
import numpy as np
import pytest

from openjij_amd import capi
from oracle import oracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu

PREC = {"fp64": capi.PREC_FP64, "fp32": capi.PREC_FP32, "int32": capi.PREC_INT32}


def sched(L, bmin=0.1, bmax=5.0, mcs=1):
    return H.geometric_betas(bmin, bmax, L), np.full(L, mcs, dtype=np.int64)


def run_dense(E_upper, prec, R, beta, mcs, seed, offset=0, init_seed=1):
    n = E_upper.shape[0] - 1
    init = H.random_pm1(R, n, init_seed)
    with capi.Context.dense(E_upper, R, 1, 0.0, PREC[prec]) as ctx:
        ctx.set_spins(init)
        ctx.run(beta, mcs, seed, offset)
        got = ctx.get_spins()
        info = ctx.info()
        energies = ctx.energies()
    want = orc.replay_dense(prec, H.symmetric_ext(E_upper), init, seed, offset,
                            orc.expand_classical(beta, mcs))
    return init, got, want, info, energies


@pytest.mark.parametrize("n", [5, 63, 64, 65, 130, 256, 700])
@pytest.mark.parametrize("prec", ["fp64", "int32"])
def test_dense_integer_couplings_bitexact(n, prec):
    E = H.int_dense(n, 1000 + n)
    beta, mcs = sched(40, 0.05, 2.0)
    init, got, want, info, en = run_dense(E, prec, 6, beta, mcs, seed=12345 + n)
    assert info["kernel"] == capi.KERNEL_DENSE_WAVE
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(en, orc.energy_exact(E, got))


@pytest.mark.parametrize("n", [17, 64, 200, 1024])
def test_dense_real_couplings_fp64_bitexact(n):
    E = H.sk_gauss(n, 2000 + n, field=True)
    beta, mcs = sched(30, 0.1, 8.0)
    _, got, want, info, en = run_dense(E, "fp64", 4, beta, mcs, seed=99)
    np.testing.assert_array_equal(got, want)
    ex = orc.energy_exact(E, got)
    assert np.all(np.abs(en - ex) <= 1e-12 * (1 + np.abs(ex)))


@pytest.mark.parametrize("n", [33, 256, 2048])
def test_dense_fp32_bitexact_without_refresh(n):
    E = H.sk_gauss(n, 3000 + n, field=True)
    beta, mcs = sched(20, 0.1, 8.0)
    _, got, want, info, _ = run_dense(E, "fp32", 3, beta, mcs, seed=7)
    assert info["coupling_bytes"] == 4
    np.testing.assert_array_equal(got, want)


def test_dense_fp64_and_int32_paths_agree():
    E = H.sk_pm1(300, 5)
    beta, mcs = sched(25, 0.1, 3.0)
    _, a, _, _, _ = run_dense(E, "fp64", 8, beta, mcs, seed=5)
    _, b, _, info, _ = run_dense(E, "int32", 8, beta, mcs, seed=5)
    assert info["coupling_bytes"] == 1
    np.testing.assert_array_equal(a, b)


def test_split_run_equals_single_run():
    E = H.int_dense(150, 77)
    beta, mcs = sched(30)
    init = H.random_pm1(4, 150, 3)
    with capi.Context.dense(E, 4) as c1, capi.Context.dense(E, 4) as c2:
        c1.set_spins(init)
        c1.run(beta, mcs, 42, 0)
        c2.set_spins(init)
        c2.run(beta[:13], mcs[:13], 42, 0)
        c2.run(beta[13:], mcs[13:], 42, 13)
        np.testing.assert_array_equal(c1.get_spins(), c2.get_spins())


@pytest.mark.parametrize("prec,n", [("fp64", 200), ("fp32", 300), ("fp32", 2048)])
def test_split_run_real_couplings_carries_fields(prec, n):
    """Real couplings: a run in three pieces (and one sweep per call) keeps
    the carried fields between calls, so it follows the single run's
    trajectory bit for bit, as the reference system keeps dE across calls."""
    E = H.sk_gauss(n, 0x5B1 + n, field=True)
    beta, mcs = sched(12, 0.1, 6.0)
    init = H.random_pm1(5, n, 8)
    with capi.Context.dense(E, 5, 1, 0.0, PREC[prec]) as c1, \
            capi.Context.dense(E, 5, 1, 0.0, PREC[prec]) as c2, \
            capi.Context.dense(E, 5, 1, 0.0, PREC[prec]) as c3:
        for c in (c1, c2, c3):
            c.set_spins(init)
        c1.run(beta, mcs, 42, 0)
        for a, b in ((0, 5), (5, 6), (6, 12)):
            c2.run(beta[a:b], mcs[a:b], 42, a)
        for t in range(12):
            c3.run(beta[t:t + 1], mcs[t:t + 1], 42, t)
        want = orc.replay_dense(prec, H.symmetric_ext(E), init, 42, 0,
                                orc.expand_classical(beta, mcs))
        np.testing.assert_array_equal(c1.get_spins(), want)
        np.testing.assert_array_equal(c2.get_spins(), want)
        np.testing.assert_array_equal(c3.get_spins(), want)
        # new spins: the fields are rebuilt from them
        c2.set_spins(init)
        c2.run(beta, mcs, 42, 0)
        np.testing.assert_array_equal(c2.get_spins(), want)


def test_config_a_sk256_pm1_replay():
    """Dense SK N=256, J = +-1, 64 reads, 1000 sweeps, beta 0.1 -> 10."""
    E = H.sk_pm1(256, 0xA001)
    beta, mcs = sched(1000, 0.1, 10.0)
    for prec in ("fp64", "int32"):
        _, got, want, info, en = run_dense(E, prec, 64, beta, mcs, seed=0xA001)
        np.testing.assert_array_equal(got, want)
        assert info["last_attempts"] == 64 * 1000 * 256
        np.testing.assert_array_equal(en, orc.energy_exact(E, got))


@pytest.mark.parametrize("n,prec,R", [(256, "int32", 64), (256, "int32", 300), (200, "int32", 5),
                                      (64, "fp64", 9), (64, "fp32", 70), (40, "fp64", 3)])
@pytest.mark.parametrize("wpb", [0, 1, 3, 4])
def test_dense_wave_rows_in_lds_bitexact(n, prec, R, wpb):
    """The wave kernel with the coupling rows staged in LDS (rows of up to
    64 KiB in total), for every replicas-per-workgroup choice, against the
    replay and against the same context reading its rows from L2."""
    E = H.sk_pm1(n, 0x1D5 + n) if prec == "int32" else H.sk_gauss(n, 0x1D5 + n, field=True)
    beta, mcs = sched(60, 0.1, 6.0)
    init = H.random_pm1(R, n, 21)
    out = {}
    for lds in (1, 0):
        with capi.Context.dense(E, R, 1, 0.0, PREC[prec]) as ctx:
            ctx.set_option("wave_helpers", 0)  # the form without threshold helpers
            ctx.set_option("lds_rows", lds)
            ctx.set_option("lds_rows_wpb", wpb)
            ctx.set_spins(init)
            ctx.run(beta, mcs, 0x1D5, 0)
            info = ctx.info()
            assert info["kernel"] == capi.KERNEL_DENSE_WAVE
            rb = ((n + 63) // 64 + 3) // 4 * 256 if prec == "int32" else \
                (n + 63) // 64 * 64 * (8 if prec == "fp64" else 4)
            assert info["lds_bytes"] == (n * rb if lds else 0)
            out[lds] = ctx.get_spins()
    want = orc.replay_dense(prec, H.symmetric_ext(E), init, 0x1D5, 0,
                            orc.expand_classical(beta, mcs))
    np.testing.assert_array_equal(out[1], want)
    np.testing.assert_array_equal(out[0], want)


@pytest.mark.parametrize("n,prec,R,L", [(256, "int32", 64, 37), (200, "int32", 3, 9),
                                        (500, "int32", 5, 6), (64, "fp64", 9, 13),
                                        (130, "fp32", 7, 5), (40, "fp64", 2, 2), (512, "fp32", 3, 4)])
@pytest.mark.parametrize("helpers", [1, 2, 3])
def test_dense_wave_threshold_helpers_bitexact(n, prec, R, L, helpers):
    """One replica per workgroup, thresholds made one phase ahead by helper
    waves (phases of 4 sweeps, schedules of any length): the replay's
    trajectory, and a run split in two carries its fields as before."""
    E = H.sk_pm1(n, 0x4E1 + n) if prec == "int32" else H.sk_gauss(n, 0x4E1 + n, field=True)
    beta, mcs = sched(L, 0.1, 6.0)
    init = H.random_pm1(R, n, 23)
    want = orc.replay_dense(prec, H.symmetric_ext(E), init, 0x4E1, 0,
                            orc.expand_classical(beta, mcs))
    with capi.Context.dense(E, R, 1, 0.0, PREC[prec]) as ctx:
        assert ctx.info()["kernel"] == capi.KERNEL_DENSE_WAVE
        ctx.set_option("wave_helpers", helpers)
        ctx.set_spins(init)
        ctx.run(beta, mcs, 0x4E1, 0)
        info = ctx.info()
        assert info["waves_per_replica"] == 1 + helpers
        np.testing.assert_array_equal(ctx.get_spins(), want)
        ctx.set_spins(init)
        h = L // 2
        ctx.run(beta[:h], mcs[:h], 0x4E1, 0)
        ctx.run(beta[h:], mcs[h:], 0x4E1, h)
        np.testing.assert_array_equal(ctx.get_spins(), want)
        np.testing.assert_array_equal(ctx.energies(), orc.energy_exact(E, want)) \
            if prec == "int32" else None


# --- CSR --------------------------------------------------------------------
@pytest.mark.parametrize("L,R,sweeps", [(3, 4, 30), (6, 16, 20), (16, 4, 6), (4, 600, 8), (3, 2100, 6)])
def test_csr_ea3d_int_bitexact(L, R, sweeps):
    n, rp, ci, v = H.ea3d_csr(L, 0xB002 + L)
    beta, mcs = sched(sweeps, 0.1, 5.0)
    init = H.random_pm1(R, n, L)
    with capi.Context.csr(n, rp, ci, v, R, precision=capi.PREC_INT32) as ctx:
        assert ctx.info()["kernel"] == capi.KERNEL_CSR_LEVEL
        ctx.set_spins(init)
        ctx.run(beta, mcs, 31337, 0)
        got = ctx.get_spins()
        en = ctx.energies()
    want = orc.replay_csr("int32", rp, ci, v, n, init, 31337, 0, orc.expand_classical(beta, mcs))
    np.testing.assert_array_equal(got, want)
    for r in range(R):
        assert en[r] == orc.energy_exact_csr(rp, ci, v, n, got[r])


def gauss_sparse_csr(n, deg, seed):
    """Random sparse graph, Gaussian couplings and fields (not FP32-exact),
    as the full symmetric (n+1)-row CSR with corner 1."""
    rng = np.random.default_rng(seed)
    pairs = set()
    while len(pairs) < n * deg // 2:
        i, j = rng.integers(0, n, 2)
        if i != j:
            pairs.add((min(i, j), max(i, j)))
    rows = [[] for _ in range(n + 1)]
    for i, j in sorted(pairs):
        v = rng.normal() / 3.0
        rows[i].append((j, v))
        rows[j].append((i, v))
    for i in range(n):
        hv = rng.normal() / 5.0
        rows[i].append((n, hv))
        rows[n].append((i, hv))
    rows[n].append((n, 1.0))
    rp, ci, va = [0], [], []
    for r in rows:
        for j, v in sorted(r):
            ci.append(j)
            va.append(v)
        rp.append(len(ci))
    return np.array(rp, np.int64), np.array(ci, np.int64), np.array(va, float)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "int32"])
def test_csr_energies_from_level_order_arrays(prec):
    """Energies of sparse systems come from the CSR kernel's own ELL arrays
    with FP64 values: exact energies of the original couplings for every
    precision of the sweep (integer couplings equal, real ones to 1e-12)."""
    if prec == "int32":
        n, rp, ci, v = H.ea3d_csr(9, 0xE9)
    else:
        n = 3000
        rp, ci, v = gauss_sparse_csr(n, 5, 0xE5)
    R = 7
    beta, mcs = sched(10, 0.1, 3.0)
    init = H.random_pm1(R, n, 5)
    with capi.Context.csr(n, rp, ci, v, R, precision=PREC[prec]) as ctx:
        assert ctx.info()["kernel"] == capi.KERNEL_CSR_LEVEL
        ctx.set_spins(init)
        first = ctx.energies()
        ctx.run(beta, mcs, 99, 0)
        got = ctx.get_spins()
        en = ctx.energies()
    for r in range(R):
        for e, s in ((first[r], init[r]), (en[r], got[r])):
            ex = orc.energy_exact_csr(rp, ci, v, n, s)
            if prec == "int32":
                assert e == ex
            else:
                assert abs(e - ex) <= 1e-12 * (1 + abs(ex)), (r, e, ex)


@pytest.mark.parametrize("waves", [1, 2, 3, 4, 8])
def test_csr_waves_option_bitexact(waves):
    n, rp, ci, v = H.ea3d_csr(5, 0xB0C5)
    beta, mcs = sched(12, 0.1, 4.0)
    init = H.random_pm1(37, n, 9)
    with capi.Context.csr(n, rp, ci, v, 37, precision=capi.PREC_FP32) as ctx:
        ctx.set_option("csr_waves", waves)
        assert ctx.info()["waves_per_replica"] == waves
        ctx.set_spins(init)
        ctx.run(beta, mcs, 606, 3)
        got = ctx.get_spins()
    want = orc.replay_csr("fp32", rp, ci, v, n, init, 606, 3, orc.expand_classical(beta, mcs))
    np.testing.assert_array_equal(got, want)


def chain_csr(n, seed):
    """Open 1D chain with integer couplings and fields: level(i) = i, so a
    sweep has n levels (the most level pointers a graph can have)."""
    w = H.splitmix64_np(seed, 2 * n)
    mag = (w[:n] >> np.uint64(61)).astype(np.int64) + 1          # never zero: one chain
    J = np.where((w[:n] >> np.uint64(60)) & np.uint64(1), mag, -mag).astype(float)
    h = ((w[n:] >> np.uint64(61)).astype(np.int64) - 4).astype(float)
    rowptr, col, val = [0], [], []
    for i in range(n):
        for j, v in ((i - 1, J[i - 1] if i > 0 else 0.0), (i + 1, J[i] if i < n - 1 else 0.0),
                     (n, h[i])):
            if j >= 0 and v != 0:
                col.append(j)
                val.append(v)
        rowptr.append(len(col))
    for i in range(n):
        if h[i] != 0:
            col.append(i)
            val.append(h[i])
    col.append(n)
    val.append(1.0)
    rowptr.append(len(col))
    return np.array(rowptr, np.int64), np.array(col, np.int64), np.array(val, float)


@pytest.mark.parametrize("n,waves,prec", [(2000, 4, "int32"), (13000, 4, "int32"),
                                          (13000, 1, "fp64")])
def test_csr_chain_many_levels(n, waves, prec):
    """n levels per sweep: the workgroup form keeps the level pointers in LDS
    when they fit beside its thresholds and spins (9 bytes per spin with INT32
    thresholds, 13 with FP64 ones), the wave-per-replica form takes the rest."""
    rp, ci, v = chain_csr(n, 0xC4A1 + n)
    beta, mcs = sched(3, 0.3, 2.0)
    init = H.random_pm1(4, n, 21)
    with capi.Context.csr(n, rp, ci, v, 4, precision=PREC[prec]) as ctx:
        assert ctx.info()["waves_per_replica"] == waves
        ctx.set_spins(init)
        ctx.run(beta, mcs, 55, 0)
        got = ctx.get_spins()
    want = orc.replay_csr(prec, rp, ci, v, n, init, 55, 0, orc.expand_classical(beta, mcs))
    np.testing.assert_array_equal(got, want)


def int_sparse_csr(n, deg, vmax, seed):
    """Random sparse graph with integer couplings and fields in -vmax..vmax
    (both ends always present), as the full symmetric (n+1)-row CSR."""
    rng = np.random.default_rng(seed)
    pairs = set()
    while len(pairs) < n * deg // 2:
        i, j = rng.integers(0, n, 2)
        if i != j:
            pairs.add((min(i, j), max(i, j)))
    rows = [[] for _ in range(n + 1)]
    for k, (i, j) in enumerate(sorted(pairs)):
        v = float(vmax if k == 0 else -vmax if k == 1 else rng.integers(-vmax, vmax + 1))
        if v != 0:
            rows[i].append((j, v))
            rows[j].append((i, v))
    for i in range(n):
        hv = float(rng.integers(-vmax, vmax + 1))
        if hv != 0:
            rows[i].append((n, hv))
            rows[n].append((i, hv))
    rows[n].append((n, 1.0))
    rp, ci, va = [0], [], []
    for r in rows:
        for j, v in sorted(r):
            ci.append(j)
            va.append(v)
        rp.append(len(ci))
    return np.array(rp, np.int64), np.array(ci, np.int64), np.array(va, float)


@pytest.mark.parametrize("deg,vmax,waves", [(2, 127, 0), (2, 127, 4), (5, 127, 2),
                                            (5, 127, 1), (12, 127, 0), (12, 127, 1),
                                            (12, 127, 8), (30, 100, 0), (5, 128, 4),
                                            (12, 1000, 0)])
def test_csr_int_couplings_packed_and_plain_bitexact(deg, vmax, waves):
    """INT32 sparse sweeps with couplings within int8 form fields by 4-way
    int8 dot products (csr_pack8 = 1, the default); wider couplings, and
    csr_pack8 = 0, add int32 values one neighbour at a time.  Both equal the
    oracle's sequential sweep at every degree width (8, 16, 32, 64)."""
    n = 700
    rp, ci, v = int_sparse_csr(n, deg, vmax, 0x1B8 + deg + vmax)
    beta, mcs = sched(8, 0.01, 0.5)
    R = 6
    init = H.random_pm1(R, n, deg)
    want = orc.replay_csr("int32", rp, ci, v, n, init, 808, 2, orc.expand_classical(beta, mcs))
    for pack in (1, 0):
        with capi.Context.csr(n, rp, ci, v, R, precision=capi.PREC_INT32) as ctx:
            assert ctx.info()["kernel"] == capi.KERNEL_CSR_LEVEL
            ctx.set_option("csr_pack8", pack)
            if waves:
                ctx.set_option("csr_waves", waves)
            ctx.set_spins(init)
            ctx.run(beta, mcs, 808, 2)
            got = ctx.get_spins()
        np.testing.assert_array_equal(got, want, err_msg=f"csr_pack8={pack}")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("deg,big_h", [(5, False), (5, True), (12, False)])
def test_csr_real_fields_with_int8_couplings(prec, deg, big_h):
    """FP32 / FP64 fields over integer couplings within int8 and integer h
    take the int8 dot-product form (csr_pack8 = 1, the default); an h of
    2^21 keeps the field-type values (an FP32 partial sum could then round).
    Both forms equal the oracle's sequential sweep in the field type."""
    n = 700
    rp, ci, v = int_sparse_csr(n, deg, 127, 0xF8 + deg)
    if big_h:  # the field of spin 3: the entries (3, n) and (n, 3)
        v = v.copy()
        v[(np.repeat(np.arange(n + 1), np.diff(rp)) == 3) & (ci == n)] = 2.0 ** 21 + 1
        v[(np.repeat(np.arange(n + 1), np.diff(rp)) == n) & (ci == 3)] = 2.0 ** 21 + 1
    beta, mcs = sched(8, 0.01, 0.5)
    R = 6
    init = H.random_pm1(R, n, deg + 1)
    want = orc.replay_csr(prec, rp, ci, v, n, init, 909, 0, orc.expand_classical(beta, mcs))
    for pack in (1, 0):
        with capi.Context.csr(n, rp, ci, v, R, precision=PREC[prec]) as ctx:
            assert ctx.info()["kernel"] == capi.KERNEL_CSR_LEVEL
            ctx.set_option("csr_pack8", pack)
            ctx.set_spins(init)
            ctx.run(beta, mcs, 909, 0)
            got = ctx.get_spins()
        np.testing.assert_array_equal(got, want, err_msg=f"csr_pack8={pack}")


@pytest.mark.parametrize("prec,waves", [("fp32", 4), ("fp32", 2), ("fp32", 1), ("fp64", 0)])
def test_csr_degree_32_real_couplings(prec, waves):
    """Neighbour lists padded to 32: FP32 fields run on the workgroup form
    (two passes' records in registers), FP64 fields on the wave-per-replica
    kernel; Gaussian couplings, bit for bit against the oracle."""
    n = 900
    rp, ci, v = gauss_sparse_csr(n, 12, 0xD32)
    assert 16 < np.diff(rp)[:n].max() - 1 <= 32
    beta, mcs = sched(10, 0.1, 3.0)
    R = 9
    init = H.random_pm1(R, n, 32)
    with capi.Context.csr(n, rp, ci, v, R, precision=PREC[prec]) as ctx:
        assert ctx.info()["kernel"] == capi.KERNEL_CSR_LEVEL
        if waves:
            ctx.set_option("csr_waves", waves)
        assert ctx.info()["waves_per_replica"] == (waves if prec == "fp32" else 1)
        ctx.set_spins(init)
        ctx.run(beta, mcs, 3232, 1)
        got = ctx.get_spins()
    want = orc.replay_csr(prec, rp, ci, v, n, init, 3232, 1, orc.expand_classical(beta, mcs))
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("prec", ["int32", "fp64"])
@pytest.mark.parametrize("rule", [0, 1])
@pytest.mark.parametrize("waves", [4, 1])
def test_csr_int_thresholds_at_extreme_betas(rule, waves, prec):
    """The workgroup form holds INT32 thresholds as ceil(log-factor(U) / -beta)
    in int32: single-precision logarithms decide unless the bracket straddles
    an integer, and values beyond int32 clamp.  Betas from 1e-9 (thresholds
    ~1e10, clamped) to 1e9, with many near-integer thresholds at beta = 1 and
    0.5, replay the oracle's double comparisons bit for bit, for the
    Metropolis and heat-bath rules; FP64 fields over these int8 couplings
    take the same integer thresholds."""
    n, rp, ci, v = H.ea3d_csr(6, 0x7E5 + rule)
    v = v * 3.0
    beta = np.array([1e-9, 1e-6, 0.001, 0.25, 0.5, 1.0, 1.0, 2.0, 1e3, 1e9])
    mcs = np.full(len(beta), 2, np.int64)
    R = 40
    init = H.random_pm1(R, n, 77)
    with capi.Context.csr(n, rp, ci, v, R, precision=PREC[prec]) as ctx:
        ctx.set_option("csr_waves", waves)
        ctx.set_option("update_method", rule)
        ctx.set_spins(init)
        ctx.run(beta, mcs, 4321, 0)
        got = ctx.get_spins()
    want = orc.replay_csr(prec, rp, ci, v, n, init, 4321, 0,
                          orc.expand_classical(beta, mcs), rule=rule)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_csr_real_couplings_bitexact(prec):
    E = H.sk_gauss(90, 4242, field=True)
    # sparsify: keep ~8 neighbours per spin
    mask = (H.splitmix64_np(5, 91 * 91).reshape(91, 91) >> np.uint64(60)) == 0
    E = np.where(mask | (np.arange(91)[None, :] == 90), E, 0.0)
    S = H.symmetric_ext(E)
    rp, ci, v = orc.dense_to_csr(S)
    beta, mcs = sched(25, 0.1, 6.0)
    init = H.random_pm1(5, 90, 8)
    with capi.Context.csr(90, rp, ci, v, 5, precision=PREC[prec]) as ctx:
        assert ctx.info()["kernel"] == capi.KERNEL_CSR_LEVEL
        ctx.set_spins(init)
        ctx.run(beta, mcs, 11, 0)
        got = ctx.get_spins()
    want = orc.replay_csr(prec, rp, ci, v, 90, init, 11, 0, orc.expand_classical(beta, mcs))
    np.testing.assert_array_equal(got, want)


def test_csr_dense_like_graph_runs_on_dense_kernel():
    E = H.symmetric_ext(H.int_dense(200, 6, density=0.9))
    rp, ci, v = orc.dense_to_csr(E)
    beta, mcs = sched(10)
    init = H.random_pm1(3, 200, 4)
    with capi.Context.csr(200, rp, ci, v, 3) as ctx:
        assert ctx.info()["kernel"] == capi.KERNEL_DENSE_WAVE
        ctx.set_spins(init)
        ctx.run(beta, mcs, 2, 0)
        got = ctx.get_spins()
    want = orc.replay_dense("int32", E, init, 2, 0, orc.expand_classical(beta, mcs))
    np.testing.assert_array_equal(got, want)


# --- Trotter ------------------------------------------------------------------
def sqa_lds_slots(n, M, prec, budget):
    """Slots of every coupling row the dense Trotter kernel keeps in LDS."""
    NS = 1 << max(0, int(np.ceil(np.log2((n + 63) // 64))))
    fb = 8 if prec == "fp64" else 4
    for sl in (4, 2, 1):
        if sl <= NS and NS <= 8 and sl * 64 * NS * 64 * fb <= 160 * 1024 and \
                M * NS * 8 + sl * n * 64 * fb <= budget:
            return NS, sl
    return NS, 0


@pytest.mark.parametrize("n,M,prec", [(32, 2, "fp64"), (32, 3, "fp64"), (40, 4, "fp32"),
                                      (70, 5, "fp64"), (64, 8, "fp32"), (130, 33, "fp32"),
                                      (256, 4, "fp32"), (200, 7, "fp32"), (256, 6, "fp64"),
                                      (500, 5, "fp32")])
@pytest.mark.parametrize("rows_lds", [1, 0])
def test_sqa_dense_bitexact(n, M, prec, rows_lds):
    """rows_lds 1 (default): the first slots of every coupling row read from
    LDS, the others from L2; 0: whole rows from L2."""
    E = H.sk_gauss(n, 5000 + n + M, field=True)
    L = 20
    s = np.linspace(0.0, 1.0, L)
    beta = np.full(L, 4.0)
    mcs = np.ones(L, dtype=np.int64)
    gamma = 2.0
    R = 3
    init = np.repeat(H.random_pm1(R, n, M)[:, None, :], M, axis=1)
    with capi.Context.dense(E, R, M, gamma, PREC[prec]) as ctx:
        assert ctx.info()["kernel"] == capi.KERNEL_SQA_WAVE
        ctx.set_option("sqa_rows_lds", rows_lds)
        ctx.set_spins(init)
        ctx.run(beta, mcs, 1234, 0, s=s)
        got = ctx.get_spins()
        sol = ctx.solution()
        en = ctx.energies()
        info = ctx.info()
    NS, sl = sqa_lds_slots(n, M, prec, info["lds_budget"])
    fb = 8 if prec == "fp64" else 4
    assert info["lds_bytes"] == M * NS * 8 + (sl * n * 64 * fb if rows_lds else 0)
    a, B = orc.expand_transverse(beta, s, mcs, gamma, M)
    want = orc.replay_sqa_dense(prec, H.symmetric_ext(E), init, 1234, 0, a, B)
    np.testing.assert_array_equal(got, want)
    for r in range(R):
        ex = orc.energy_exact(E, got[r])
        assert np.all(np.abs(en[r] - ex) <= 1e-12 * (1 + np.abs(ex)))
        assert list(sol[r]) == list(got[r, int(np.argmin(en[r]))])


@pytest.mark.parametrize("n,M,prec", [(2049, 3, "fp32"), (3000, 8, "fp64"), (4096, 7, "fp64"),
                                      (4096, 16, "fp32"), (3500, 10, "fp32")])
def test_sqa_dense_wide_bitexact(n, M, prec):
    """2049-4096 spins: 64 slots per slice, one slice pair per wave (FP64) or
    up to two (FP32)."""
    print(f"dense Trotter wide: n={n} M={M} {prec}", flush=True)
    E = H.sk_gauss(n, 6000 + n + M, field=True)
    L = 4
    s = np.linspace(0.2, 0.8, L)
    beta = np.full(L, 3.0)
    mcs = np.ones(L, dtype=np.int64)
    gamma = 1.0
    R = 2
    init = np.stack([H.random_pm1(M, n, 7 + r) for r in range(R)])
    with capi.Context.dense(E, R, M, gamma, PREC[prec]) as ctx:
        assert ctx.info()["kernel"] == capi.KERNEL_SQA_WAVE
        ctx.set_spins(init)
        ctx.run(beta, mcs, 0x3E3E, 2, s=s)
        got = ctx.get_spins()
        en = ctx.energies()
    a, B = orc.expand_transverse(beta, s, mcs, gamma, M)
    want = orc.replay_sqa_dense(prec, H.symmetric_ext(E), init, 0x3E3E, 2, a, B)
    np.testing.assert_array_equal(got, want)
    for r in range(R):
        ex = orc.energy_exact(E, got[r])
        assert np.all(np.abs(en[r] - ex) <= 1e-12 * (1 + np.abs(ex)))


def test_sqa_dense_wide_limits():
    """Past the registers of the 64-slot Trotter kernel (FP64 above 8
    slices, FP32 above 16, any n above 4096) the context is refused."""
    for n, M, prec in ((3000, 9, "fp64"), (3000, 17, "fp32"), (4097, 2, "fp32")):
        with pytest.raises(capi.OjmiError):
            capi.Context.dense(H.sk_gauss(n, 7), 2, M, 1.0, PREC[prec])


def test_sqa_config_c_shape_replay():
    """SK N=256 Gaussian, M=32, Gamma=5, s linear to 4.99/5 (a short slice of
    the 1000-sweep schedule), fp32 and fp64."""
    n, M, R = 256, 32, 2
    E = H.sk_gauss(n, 0xC003)
    L = 8
    s = (4.99 / 5) * np.arange(L) / 999.0
    beta = np.full(L, 10.0)
    mcs = np.ones(L, dtype=np.int64)
    init = np.repeat(H.random_pm1(R, n, 3)[:, None, :], M, axis=1)
    a, B = orc.expand_transverse(beta, s, mcs, 5.0, M)
    for prec, rows_lds in (("fp32", 1), ("fp32", 0), ("fp64", 1)):
        with capi.Context.dense(E, R, M, 5.0, PREC[prec]) as ctx:
            ctx.set_option("sqa_rows_lds", rows_lds)
            ctx.set_spins(init)
            ctx.run(beta, mcs, 0xC003, 0, s=s)
            got = ctx.get_spins()
        want = orc.replay_sqa_dense(prec, H.symmetric_ext(E), init, 0xC003, 0, a, B)
        np.testing.assert_array_equal(got, want)


# --- ground states and statistics ---------------------------------------------
def test_reference_fixture_ground_states(fixtures):
    """SingleSpinFlip.FindTrueGroundState on the five reference fixtures with
    the reference test schedule (beta 0.1 -> 100, 100 levels x 100 sweeps):
    the GPU's ground-state rate over 256 reads agrees with the reference
    restatement's over 256 reads (two-proportion z-test, |z| < 4), and every
    fixture's ground state is found."""
    c = fixtures["classical_schedule"]
    beta = H.geometric_betas(c["beta_min"], c["beta_max"], c["num_call_updater"])
    mcs = np.full(len(beta), c["one_mc_step"], dtype=np.int64)
    R = 256
    for k, case in fixtures["testcases"].items():
        E = np.array(case["J_ext_upper"], dtype=float)
        init = H.random_pm1(R, 8, int(k))
        with capi.Context.dense(E, R) as ctx:
            ctx.set_spins(init)
            ctx.run(beta, mcs, 5678, 0)
            got = ctx.get_spins()
        ref = orc.ref_sa_dense_batch(H.symmetric_ext(E), init,
                                     np.arange(R, dtype=np.uint32) + 1, beta, mcs)
        p_gpu = np.mean([list(s) == case["ground_state"] for s in got])
        p_ref = np.mean([list(s) == case["ground_state"] for s in ref])
        pool = (p_gpu + p_ref) / 2
        se = np.sqrt(max(pool * (1 - pool), 1e-12) * 2 / R)
        print(f"fixture {k}: ground-state rate GPU {p_gpu:.3f} reference {p_ref:.3f}")
        assert p_gpu > 0.5, (k, p_gpu)
        assert abs(p_gpu - p_ref) / se < 4.0, (k, p_gpu, p_ref)


@pytest.mark.parametrize("n", [12, 18, 22])
def test_brute_force_ground_state_found(n):
    E = H.sk_gauss(n, 7000 + n, field=True)
    e0, s0 = orc.brute_force(E)
    beta, mcs = sched(300, 0.1, 10.0)
    with capi.Context.dense(E, 64) as ctx:
        ctx.set_spins(H.random_pm1(64, n, 2))
        ctx.run(beta, mcs, 3, 0)
        en = ctx.energies()
    assert en.min() == pytest.approx(e0, abs=1e-9)


def test_energy_distribution_matches_reference_ks():
    """Two-sample KS between the GPU sampler and the reference restatement
    (Xorshift + libstdc++ uniforms) at the same schedule: SK N=64 +-1, 1000
    reads each, 60 sweeps beta 0.1 -> 2.  Threshold p > 1e-3."""
    n, R = 64, 1000
    E = H.sk_pm1(n, 0x5EED)
    beta, mcs = sched(60, 0.1, 2.0)
    init = H.random_pm1(R, n, 77)
    with capi.Context.dense(E, R, precision=capi.PREC_FP64) as ctx:
        ctx.set_spins(init)
        ctx.run(beta, mcs, 2024, 0)
        e_gpu = ctx.energies()
    ref = orc.ref_sa_dense_batch(H.symmetric_ext(E), init, np.arange(R, dtype=np.uint32) + 1,
                                 beta, mcs)
    e_ref = orc.energy_exact(E, ref)
    p = H.ks_2samp_pvalue(e_gpu, e_ref)
    print(f"KS p-value GPU vs reference restatement: {p:.4f}")
    assert p > 1e-3


# --- workgroup-shared dense kernel ---------------------------------------------
@pytest.mark.parametrize("colown", [0, 1])
@pytest.mark.parametrize("n,R,G,prec", [(64, 8, 8, "fp32"), (100, 13, 4, "fp32"),
                                        (257, 40, 8, "fp32"), (1000, 33, 16, "fp32"),
                                        (300, 24, 8, "int32"), (129, 10, 4, "int32"),
                                        (4096, 16, 8, "fp32"), (600, 11, 8, "fp32"),
                                        (520, 9, 4, "int32"), (2048, 20, 4, "fp32")])
def test_shared_kernel_bitexact(n, R, G, prec, colown):
    """Replicas of a workgroup share accepted coupling rows (rows staged in LDS,
    or column-owned fields with each row read once per workgroup); the
    trajectory is still the sequential one (oracle replay, bit for bit)."""
    E = H.sk_gauss(n, 9000 + n, field=True) if prec == "fp32" else H.int_dense(n, 9000 + n)
    beta, mcs = sched(12 if n >= 4096 else 25, 0.1, 6.0)
    init = H.random_pm1(R, n, 5)
    with capi.Context.dense(E, R, 1, 0.0, PREC[prec]) as ctx:
        ctx.set_option("group", G)
        ctx.set_option("colown", colown)
        slots = 1 << max(0, (n + 63) // 64 - 1).bit_length()
        owned = colown and G in (4, 8) and slots >= 8 and slots % G == 0
        assert ctx.info()["kernel"] == (capi.KERNEL_DENSE_COLOWN if owned
                                        else capi.KERNEL_DENSE_SHARED)
        ctx.set_spins(init)
        ctx.run(beta, mcs, 777, 4)
        got = ctx.get_spins()
    want = orc.replay_dense(prec, H.symmetric_ext(E), init, 777, 4,
                            orc.expand_classical(beta, mcs))
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("colown", [0, 1])
@pytest.mark.parametrize("window", [0, 1, 3])
def test_shared_kernel_xcd_pacing_does_not_change_results(window, colown):
    n, R = 512, 256  # 32 workgroups of 8: pacing active when window > 0
    E = H.sk_gauss(n, 31)
    beta, mcs = sched(10, 0.1, 6.0)
    init = H.random_pm1(R, n, 6)
    with capi.Context.dense(E, R, 1, 0.0, capi.PREC_FP32) as ctx:
        ctx.set_option("group", 8)
        ctx.set_option("colown", colown)
        ctx.set_option("xcd_window", window)
        ctx.set_spins(init)
        ctx.run(beta, mcs, 99, 0)
        got = ctx.get_spins()
    want = orc.replay_dense("fp32", H.symmetric_ext(E), init, 99, 0,
                            orc.expand_classical(beta, mcs))
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("n,prec", [(4097, "int32"), (5000, "fp64"), (4500, "fp32"),
                                    (6000, "int32")])
def test_dense_large_n_fields_in_lds_bitexact(n, prec):
    """N > 4096: one workgroup per replica with its fields in LDS; the same
    sequential trajectory (oracle replay, bit for bit) and exact energies."""
    E = H.int_dense(n, 4000 + n) if prec == "int32" else H.sk_gauss(n, 4000 + n, field=True)
    beta, mcs = sched(3, 0.2, 3.0)
    init, got, want, info, en = run_dense(E, prec, 3, beta, mcs, seed=31 + n)
    assert info["kernel"] == capi.KERNEL_DENSE_LDS
    assert not np.array_equal(got, init)
    np.testing.assert_array_equal(got, want)
    ex = orc.energy_exact(E, got)
    assert np.all(np.abs(en - ex) <= 1e-9 * (1 + np.abs(ex)))
