"""GPU parity of the heat-bath acceptance rule (option "update_method" = 1):
every classical kernel family against the oracle's replay with rule 1, bit
for bit.  The reference's rule is 1 / (1 + exp(beta dE)) > U
(include/openjij/updater/single_spin_flip.hpp:365-371)."""
import numpy as np
import pytest

from openjij_amd import capi
from oracle import oracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu

PREC = {"fp64": capi.PREC_FP64, "fp32": capi.PREC_FP32, "int32": capi.PREC_INT32}
HEAT_BATH = 1


def sched(L, bmin=0.1, bmax=5.0):
    return H.geometric_betas(bmin, bmax, L), np.ones(L, dtype=np.int64)


def run_dense(E, prec, R, L, seed, expect_kernel, options=()):
    n = E.shape[0] - 1
    beta, mcs = sched(L)
    init = H.random_pm1(R, n, seed + 1)
    with capi.Context.dense(E, R, 1, 0.0, PREC[prec]) as ctx:
        for k, v in options:
            ctx.set_option(k, v)
        ctx.set_option("update_method", HEAT_BATH)
        assert ctx.info()["kernel"] == expect_kernel
        ctx.set_spins(init)
        ctx.run(beta, mcs, seed, 2)
        got = ctx.get_spins()
        en = ctx.energies()
    want = orc.replay_dense(prec, H.symmetric_ext(E), init, seed, 2,
                            orc.expand_classical(beta, mcs), rule=1)
    np.testing.assert_array_equal(got, want)
    metro = orc.replay_dense(prec, H.symmetric_ext(E), init, seed, 2,
                             orc.expand_classical(beta, mcs), rule=0)
    assert not np.array_equal(want, metro)  # the rule changes the trajectory
    ex = orc.energy_exact(E, got)
    assert np.all(np.abs(en - ex) <= 1e-9 * (1 + np.abs(ex)))


@pytest.mark.parametrize("prec", ["fp64", "int32", "fp32"])
def test_heat_bath_dense_wave(prec):
    E = H.int_dense(150, 61) if prec == "int32" else H.sk_gauss(150, 61, field=True)
    run_dense(E, prec, 4, 15, 100, capi.KERNEL_DENSE_WAVE)


@pytest.mark.parametrize("prec", ["fp32", "int32"])
def test_heat_bath_dense_colown(prec):
    E = H.int_dense(600, 62) if prec == "int32" else H.sk_gauss(600, 62, field=True)
    run_dense(E, prec, 16, 8, 200, capi.KERNEL_DENSE_COLOWN, options=(("group", 4),))


def test_heat_bath_colown_on_a_group8_context_with_pacing():
    """Heat-bath colown kernels exist for G = 4; a G = 8 context runs them
    with G = 4 workgroups (and a pacing buffer sized for them)."""
    E = H.sk_gauss(600, 66, field=True)
    run_dense(E, "fp32", 64, 6, 500, capi.KERNEL_DENSE_COLOWN,
              options=(("group", 8), ("xcd_window", 2)))


def test_heat_bath_dense_shared():
    E = H.sk_gauss(300, 63, field=True)
    run_dense(E, "fp32", 16, 8, 300, capi.KERNEL_DENSE_SHARED,
              options=(("group", 8), ("colown", 0)))


def test_heat_bath_dense_fields_in_lds():
    E = H.int_dense(4200, 64)
    run_dense(E, "int32", 2, 2, 400, capi.KERNEL_DENSE_LDS)


@pytest.mark.parametrize("prec,waves", [("int32", 0), ("fp32", 1), ("fp64", 4)])
def test_heat_bath_csr(prec, waves):
    n, rp, ci, v = H.ea3d_csr(5, 0xB0C7)
    if prec != "int32":
        v = v * 0.731
    beta, mcs = sched(12, 0.1, 4.0)
    init = H.random_pm1(24, n, 19)
    with capi.Context.csr(n, rp, ci, v, 24, precision=PREC[prec]) as ctx:
        ctx.set_option("csr_waves", waves)
        ctx.set_option("update_method", HEAT_BATH)
        assert ctx.info()["kernel"] == capi.KERNEL_CSR_LEVEL
        ctx.set_spins(init)
        ctx.run(beta, mcs, 808, 0)
        got = ctx.get_spins()
    want = orc.replay_csr(prec, rp, ci, v, n, init, 808, 0, orc.expand_classical(beta, mcs),
                          rule=1)
    np.testing.assert_array_equal(got, want)


def test_heat_bath_batch():
    sizes = [5, 64, 130, 300]
    Es = [H.int_dense(n, 900 + n) for n in sizes]
    R, L = 3, 10
    beta, mcs = sched(L, 0.2, 3.0)
    betas = np.stack([beta * (1 + 0.1 * k) for k in range(len(sizes))])
    inits = [H.random_pm1(R, n, 50 + k) for k, n in enumerate(sizes)]
    with capi.BatchContext.create(Es, R, capi.PREC_INT32) as ctx:
        ctx.set_option("update_method", HEAT_BATH)
        ctx.set_spins(inits)
        ctx.run_batch(betas, mcs, 77, 0)
        got = ctx.get_spins()
    for k, E in enumerate(Es):
        want = orc.replay_dense("int32", H.symmetric_ext(E), inits[k], 77, 0,
                                orc.expand_classical(betas[k], mcs), word3=k, rule=1)
        np.testing.assert_array_equal(got[k], want, err_msg=f"problem {k}")


def test_heat_bath_rejected_for_trotter_systems():
    E = H.sk_gauss(20, 65)
    with capi.Context.dense(E, 2, 4, 1.0, capi.PREC_FP64) as ctx:
        with pytest.raises(capi.OjmiError):
            ctx.set_option("update_method", HEAT_BATH)
        ctx.set_option("update_method", 0)
