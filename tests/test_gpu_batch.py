"""Batch contexts (ojmi_create_batch / ojmi_run_batch) on the GPU: every
problem of a batch follows the oracle replay of its own instance bit for bit,
with the problem index as the fourth counter word, and batch energies equal
the exact energies."""
import numpy as np
import pytest

from openjij_amd import capi
from oracle import oracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu

PREC = {"fp64": capi.PREC_FP64, "fp32": capi.PREC_FP32, "int32": capi.PREC_INT32}
SIZES = [1, 2, 63, 64, 65, 129, 200, 256, 257, 300, 511, 77]


def problems(sizes, prec, seed0):
    out = []
    for k, n in enumerate(sizes):
        if prec == "int32":
            out.append(H.int_dense(n, seed0 + k))
        else:
            out.append(H.sk_gauss(n, seed0 + k, field=True))
    return out


def per_problem_betas(K, L, seed):
    rng = np.random.default_rng(seed)
    lo = rng.uniform(0.05, 0.5, K)
    hi = rng.uniform(2.0, 8.0, K)
    return np.stack([H.geometric_betas(lo[k], hi[k], L) for k in range(K)])


@pytest.mark.parametrize("prec", ["int32", "fp64", "fp32"])
def test_batch_bitexact_per_problem(prec):
    Es = problems(SIZES, prec, 700)
    R, L = 3, 12
    mcs = np.array([1, 2] * (L // 2), np.int64)
    betas = per_problem_betas(len(Es), L, 5)
    inits = [H.random_pm1(R, E.shape[0] - 1, 40 + k) for k, E in enumerate(Es)]
    with capi.BatchContext.create(Es, R, PREC[prec]) as ctx:
        info = ctx.info()
        assert info["kernel"] == capi.KERNEL_BATCH_FUSED
        assert info["num_problems"] == len(Es)
        ctx.set_spins(inits)
        ctx.run_batch(betas, mcs, 2024, 9)
        got = ctx.get_spins()
        assert info["num_spins"] == sum(SIZES)
    for k, E in enumerate(Es):
        want = orc.replay_dense(prec, H.symmetric_ext(E), inits[k], 2024, 9,
                                orc.expand_classical(betas[k], mcs), word3=k)
        np.testing.assert_array_equal(got[k], want, err_msg=f"problem {k} (n={SIZES[k]})")


@pytest.mark.parametrize("prec", ["int32", "fp64"])
def test_batch_energies_are_exact(prec):
    Es = problems(SIZES, prec, 900)
    R = 4
    with capi.BatchContext.create(Es, R, PREC[prec]) as ctx:
        ctx.set_spins([H.random_pm1(R, E.shape[0] - 1, 3 + k) for k, E in enumerate(Es)])
        ctx.run_batch(per_problem_betas(len(Es), 5, 1), np.ones(5, np.int64), 3)
        e = ctx.energies()
        spins = ctx.get_spins()
    for k, E in enumerate(Es):
        want = orc.energy_exact(H.symmetric_ext(E), spins[k])
        if prec == "int32":
            np.testing.assert_array_equal(e[k], want)
        else:
            np.testing.assert_allclose(e[k], want, rtol=1e-12, atol=1e-12)


def test_batch_common_schedule_and_split_runs():
    Es = problems([40, 300, 130], "int32", 11)
    R = 5
    beta, mcs = H.geometric_betas(0.1, 4.0, 10), np.ones(10, np.int64)
    inits = [H.random_pm1(R, E.shape[0] - 1, 8 + k) for k, E in enumerate(Es)]
    with capi.BatchContext.create(Es, R) as a, capi.BatchContext.create(Es, R) as b:
        a.set_spins(inits)
        a.run(beta, mcs, 77, 0)  # ojmi_run: one schedule for every problem
        b.set_spins(inits)
        rows = np.tile(beta, (3, 1))
        b.run_batch(rows[:, :4], mcs[:4], 77, 0)
        b.run_batch(rows[:, 4:], mcs[4:], 77, 4)
        for x, y in zip(a.get_spins(), b.get_spins()):
            np.testing.assert_array_equal(x, y)
        assert a.info()["last_attempts"] == 10 * R * sum(E.shape[0] - 1 for E in Es)


def test_batch_large_problems_replay():
    Es = problems([1024, 2048, 4096], "int32", 31)
    R, L = 2, 3
    betas = per_problem_betas(3, L, 2)
    mcs = np.ones(L, np.int64)
    inits = [H.random_pm1(R, E.shape[0] - 1, 60 + k) for k, E in enumerate(Es)]
    with capi.BatchContext.create(Es, R) as ctx:
        ctx.set_spins(inits)
        ctx.run_batch(betas, mcs, 5, 0)
        got = ctx.get_spins()
    for k, E in enumerate(Es):
        want = orc.replay_dense("int32", H.symmetric_ext(E), inits[k], 5, 0,
                                orc.expand_classical(betas[k], mcs), word3=k)
        np.testing.assert_array_equal(got[k], want)


def test_batch_argument_errors():
    Es = problems([10, 20], "int32", 1)
    with capi.BatchContext.create(Es, 2) as ctx:
        with pytest.raises(ValueError):
            ctx.run_batch(np.ones((3, 4)), np.ones(4, np.int64), 1)
        with pytest.raises(capi.OjmiError):
            ctx.run_batch(-np.ones((2, 4)), np.ones(4, np.int64), 1)
        with pytest.raises(capi.OjmiError):
            ctx.set_spins([np.zeros((2, 10), np.int8), np.ones((2, 20), np.int8)])
    with capi.Context.dense(Es[0], 2) as d:
        rc = capi.lib().ojmi_run_batch(d.h, None, None, 0, 0, 0)
        assert rc == capi.EINVAL


@pytest.mark.parametrize("mode,waves", [(0, 4), (2, 1), (2, 2), (2, 4)])
def test_batch_launch_modes_agree(mode, waves):
    """One launch of all problems largest first (each at its own slot count
    inside the widest slot capacity), the persistent launch (per-XCD
    longest-first queues, any wave may run any replica) and one launch per
    slot class give the same spins bit for bit; 200 problems so that every
    queue holds many items and stealing happens at the end."""
    ns = H.qubo_batch_sizes(200, 0xD0D)
    ns = np.minimum(ns, 700)
    Es = [H.qubo_to_scaled_ising(H.qubo_upper(int(n), 0xD0E + k))[0] for k, n in enumerate(ns)]
    R, L = 3, 4
    betas = per_problem_betas(len(Es), L, 9)
    mcs = np.ones(L, np.int64)
    inits = [H.random_pm1(R, E.shape[0] - 1, 70 + k) for k, E in enumerate(Es)]
    out = []
    kernels = {0: capi.KERNEL_BATCH_FUSED, 1: capi.KERNEL_BATCH_WAVE, 2: capi.KERNEL_BATCH_PERSIST}
    for m in (mode, 1):
        with capi.BatchContext.create(Es, R, capi.PREC_INT32) as ctx:
            ctx.set_option("batch_mode", m)
            if m == 2:
                ctx.set_option("batch_waves", waves)
            assert ctx.info()["kernel"] == kernels[m]
            ctx.set_spins(inits)
            ctx.run_batch(betas, mcs, 321, 0)
            out.append(ctx.get_spins())
            flips = ctx.info()["last_flips"]
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
    assert flips > 0
