"""The exact kernel instances bench.py times, under bit-exact parity.

* Dense SK N=4096, FP32: a 2048-replica context alternates, per 16-sweep
  segment, between the matrix-core sweep (8 replicas per workgroup; its
  instance at this shape is pinned in test_gpu_mfma_sweep.py) and the
  column-owned kernel of the autotuner's shape (4 replicas per workgroup,
  16 slots per wave chunk, as tuned on MI355X); the latter runs here for a
  few sweeps of 8 replicas against the FP32 replay.
* 4096 uneven QUBOs, 16 reads each: 128 problems drawn from the same size
  distribution, QUBO -> Ising with the exact x4 scaling, so that the
  workgroup order maps whole rounds of problems onto XCDs (non-identity
  chunking) in every slot class that has more than one round; every problem
  bit-exact against its own replay with the problem index as counter word.
"""
import numpy as np
import pytest

from openjij_amd import capi
from oracle import oracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu


def test_config_e_kernel_shape_bitexact():
    E = H.sk_gauss(4096, 0xE005)
    R, L = 8, 4
    beta = H.geometric_betas(0.1, 10.0, 500)[:L]
    mcs = np.ones(L, np.int64)
    init = H.random_pm1(R, 4096, 0xE1)
    with capi.Context.dense(E, 2048, 1, 0.0, capi.PREC_FP32) as big:
        shape = big.info()  # the launch shape config E's context runs with
        assert (shape["kernel"], shape["group"]) == (capi.KERNEL_DENSE_MFMA, 8)
        big.set_option("mfma_sweep", 0)
        colown = big.info()  # the column-owned kernel it alternates with
    assert (colown["kernel"], colown["group"]) == (capi.KERNEL_DENSE_COLOWN, 4)
    with capi.Context.dense(E, R, 1, 0.0, capi.PREC_FP32) as ctx:
        ctx.set_option("group", colown["group"])
        ctx.set_option("colown", 1)
        info = ctx.info()
        assert (info["kernel"], info["group"]) == (colown["kernel"], colown["group"])
        ctx.set_spins(init)
        ctx.run(beta, mcs, 1000, 0)
        got = ctx.get_spins()
    want = orc.replay_dense("fp32", H.symmetric_ext(E), init, 1000, 0,
                            orc.expand_classical(beta, mcs))
    assert not np.array_equal(got, init)
    np.testing.assert_array_equal(got, want)


def _qubo_batch(K, seed):
    ns = H.qubo_batch_sizes(K, seed)
    Es, betas = [], []
    for k, n in enumerate(ns):
        Q = H.qubo_upper(int(n), seed + 1 + k)
        E, _ = H.qubo_to_scaled_ising(Q)
        Es.append(E)
    return ns, Es


def test_config_d_launch_shape_bitexact():
    K, R, L = 128, 16, 3
    ns, Es = _qubo_batch(K, 0xD004)
    # slot classes of 64 spins: the XCD-chunked order is a non-identity map
    # only when a class has at least one full round (8 XCDs x 4 workgroups)
    slots = np.array([1 << max(0, (int(n) + 63) // 64 - 1).bit_length() for n in ns])
    per_class = {s: int((slots == s).sum()) for s in np.unique(slots)}
    assert max(c * R // 4 for c in per_class.values()) >= 64, per_class
    rng = np.random.default_rng(3)
    betas = np.stack([H.geometric_betas(rng.uniform(0.01, 0.05), rng.uniform(0.5, 1.2), L)
                      for _ in range(K)])
    mcs = np.ones(L, np.int64)
    inits = [H.random_pm1(R, int(n), 500 + k) for k, n in enumerate(ns)]
    with capi.BatchContext.create(Es, R, capi.PREC_INT32) as ctx:
        ctx.set_spins(inits)
        ctx.run_batch(betas, mcs, 0xD004, 0)
        got = ctx.get_spins()
        en = ctx.energies()
    for k, E in enumerate(Es):
        S = H.symmetric_ext(E)
        want = orc.replay_dense("int32", S, inits[k], 0xD004, 0,
                                orc.expand_classical(betas[k], mcs), word3=k)
        np.testing.assert_array_equal(got[k], want, err_msg=f"problem {k} (n={ns[k]})")
        np.testing.assert_array_equal(en[k], orc.energy_exact(S, got[k]))
