"""Suzuki-Trotter sweeps of sparse systems on the CSR Trotter kernel
(k_sqa_csr): every slice of a replica in LDS, the slices of one phase (even,
odd, last when M is odd) in parallel, spins of a slice in level order.

Bit-exact against the oracle's CSR Trotter replay (fields recomputed per
attempt in stored CSR order, reference include/openjij/system/
transverse_ising.hpp:426-596 under updater/single_spin_flip.hpp:128-251) for
integer and Gaussian couplings, FP64 and FP32, odd and even M, the four
neighbour-list widths, systems beyond the dense Trotter kernel's 4096 spins
and the largest n*M the LDS holds.  Energies per slice against the exact FP64
evaluation; the solution is the slice of lowest energy.
"""
import numpy as np
import pytest

from openjij_amd import capi
from oracle import oracle as orc
from tests import helpers as H
from tests.test_gpu_engine import gauss_sparse_csr

pytestmark = pytest.mark.gpu

PREC = {"fp64": capi.PREC_FP64, "fp32": capi.PREC_FP32}


def graph(kind, size, seed):
    if kind == "ea3d":
        return H.ea3d_csr(size[0], seed)
    n, deg = size
    rp, ci, va = gauss_sparse_csr(n, deg, seed)
    return (n, rp, ci, va)


def max_degree(rp, ci, n):
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    keep = (rows < n) & (ci < n) & (ci != rows)
    return int(np.bincount(rows[keep], minlength=n).max())


def trotter_schedule(L, gamma, M, bmax=4.0):
    s = np.linspace(0.05, 0.95, L)
    beta = np.full(L, bmax)
    mcs = np.ones(L, dtype=np.int64)
    a, B = orc.expand_transverse(beta, s, mcs, gamma, M)
    return beta, s, mcs, a, B


def check_energies(rp, ci, va, n, spins, en):
    R, M = spins.shape[:2]
    for r in range(R):
        for t in range(M):
            ex = orc.energy_exact_csr(rp, ci, va, n, spins[r, t])
            assert abs(en[r, t] - ex) <= 1e-12 * (1 + abs(ex))


CASES = [
    ("ea3d", (4,), 4, "fp64", 3, 12),
    ("ea3d", (5,), 5, "fp32", 4, 10),
    ("ea3d", (5,), 2, "fp64", 3, 10),
    ("gauss", (200, 5), 3, "fp64", 5, 10),
    ("gauss", (200, 5), 6, "fp32", 5, 10),
    ("gauss", (400, 12), 7, "fp32", 3, 6),
    ("gauss", (300, 12), 8, "fp64", 3, 6),
    ("gauss", (150, 30), 4, "fp64", 3, 6),
    ("gauss", (150, 30), 9, "fp32", 3, 6),
]


@pytest.mark.parametrize("kind,size,M,prec,R,L", CASES)
def test_sqa_csr_bitexact(kind, size, M, prec, R, L):
    n, rp, ci, va = graph(kind, size, 0x5C00 + M + sum(size))
    print(f"CSR Trotter: {kind} n={n} max degree {max_degree(rp, ci, n)} M={M} {prec} R={R} "
          f"sweeps={L}", flush=True)
    gamma = 1.5
    beta, s, mcs, a, B = trotter_schedule(L, gamma, M)
    init = np.stack([H.random_pm1(M, n, 17 + r) for r in range(R)])  # independent slices
    with capi.Context.csr(n, rp, ci, va, R, M, gamma, PREC[prec]) as ctx:
        info = ctx.info()
        assert info["kernel"] == capi.KERNEL_SQA_CSR and info["num_trotter"] == M
        ctx.set_spins(init)
        ctx.run(beta, mcs, 0x7777, 5, s=s)
        got = ctx.get_spins()
        en = ctx.energies()
        sol = ctx.solution()
        info = ctx.info()
    want = orc.replay_sqa_csr(prec, rp, ci, va, n, init, 0x7777, 5, a, B)
    np.testing.assert_array_equal(got, want)
    assert info["last_attempts"] == L * R * M * n
    # spins of the slices, then (when they fit) each wave's Philox words
    staged = ((n * M + 15) // 16) * 16 + min(16, (M + 1) // 2) * 16 * ((n + 3) // 4)
    assert info["lds_bytes"] in (n * M, staged)
    check_energies(rp, ci, va, n, got, en.reshape(R, M))
    for r in range(R):
        assert list(sol[r]) == list(got[r, int(np.argmin(en.reshape(R, M)[r]))])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("kind,size", [("ea3d", (6,)), ("gauss", (300, 8))])
def test_sqa_csr_records_and_word_loads_agree(prec, kind, size):
    """Neighbour lists of <= 16 load as per-position 16-byte records (the
    default; EA3D's +-1 couplings as int8 quadruples) or one word per column
    and value (sqa_csr_records = 0): the same spins, equal to the oracle's
    replay."""
    n, rp, ci, va = graph(kind, size, 0x5C0D)
    assert 8 < max_degree(rp, ci, n) <= 16 if kind == "gauss" else True
    M, R, L, gamma = 6, 5, 8, 1.5
    beta, s, mcs, a, B = trotter_schedule(L, gamma, M)
    init = np.stack([H.random_pm1(M, n, 29 + r) for r in range(R)])
    want = orc.replay_sqa_csr(prec, rp, ci, va, n, init, 0x6161, 0, a, B)
    for rec in (1, 0):
        with capi.Context.csr(n, rp, ci, va, R, M, gamma, PREC[prec]) as ctx:
            ctx.set_option("sqa_csr_records", rec)
            ctx.set_spins(init)
            ctx.run(beta, mcs, 0x6161, 0, s=s)
            np.testing.assert_array_equal(ctx.get_spins(), want, err_msg=f"records={rec}")


def test_sqa_csr_split_run_equals_one_run():
    n = 500
    rp, ci, va = gauss_sparse_csr(n, 6, 0x5D01)
    M, R, L, gamma = 5, 6, 12, 2.0
    beta, s, mcs, a, B = trotter_schedule(L, gamma, M)
    init = np.stack([H.random_pm1(M, n, 40 + r) for r in range(R)])
    with capi.Context.csr(n, rp, ci, va, R, M, gamma, capi.PREC_FP64) as ctx:
        ctx.set_spins(init)
        ctx.run(beta[:5], mcs[:5], 3, 0, s=s[:5])
        ctx.run(beta[5:], mcs[5:], 3, 5, s=s[5:])
        got = ctx.get_spins()
        flips = ctx.info()["last_flips"]
    want = orc.replay_sqa_csr("fp64", rp, ci, va, n, init, 3, 0, a, B)
    np.testing.assert_array_equal(got, want)
    assert flips > 0


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_sqa_csr_beyond_dense_limit(prec):
    """EA3D L = 17 (4913 spins, more than the dense Trotter kernel's 4096),
    M = 8."""
    n, rp, ci, va = H.ea3d_csr(17, 0x5E17)
    M, R, L, gamma = 8, 2, 3, 1.0
    beta, s, mcs, a, B = trotter_schedule(L, gamma, M, bmax=2.0)
    init = np.stack([H.random_pm1(M, n, 70 + r) for r in range(R)])
    with capi.Context.csr(n, rp, ci, va, R, M, gamma, PREC[prec]) as ctx:
        assert ctx.info()["kernel"] == capi.KERNEL_SQA_CSR
        ctx.set_spins(init)
        ctx.run(beta, mcs, 0x17, 0, s=s)
        got = ctx.get_spins()
        en = ctx.energies()
    want = orc.replay_sqa_csr(prec, rp, ci, va, n, init, 0x17, 0, a, B)
    np.testing.assert_array_equal(got, want)
    check_energies(rp, ci, va, n, got, en.reshape(R, M))


def test_sqa_csr_largest_lds_footprint():
    """n * M = 160000 bytes of slice spins: the most the kernel's LDS holds
    on a 160 KiB device, Gaussian couplings, FP32."""
    n, M = 5000, 32
    rp, ci, va = gauss_sparse_csr(n, 4, 0x5F00)
    R, L, gamma = 2, 2, 1.0
    beta, s, mcs, a, B = trotter_schedule(L, gamma, M)
    init = np.stack([H.random_pm1(M, n, 90 + r) for r in range(R)])
    with capi.Context.csr(n, rp, ci, va, R, M, gamma, capi.PREC_FP32) as ctx:
        info = ctx.info()
        if n * M > info["lds_budget"]:
            pytest.skip(f"device LDS budget {info['lds_budget']} below {n * M}")
        assert info["kernel"] == capi.KERNEL_SQA_CSR
        ctx.set_spins(init)
        ctx.run(beta, mcs, 0x60, 0, s=s)
        got = ctx.get_spins()
    want = orc.replay_sqa_csr("fp32", rp, ci, va, n, init, 0x60, 0, a, B)
    np.testing.assert_array_equal(got, want)


def test_sqa_csr_routing():
    """A node of degree > 64 sends the system to the dense Trotter kernel;
    slices that do not fit the LDS (n * M above the budget, which also
    exceeds what the dense Trotter kernel's registers hold) are refused."""
    rp, ci, va = gauss_sparse_csr(300, 80, 0x6000)
    assert max_degree(rp, ci, 300) > 64
    with capi.Context.csr(300, rp, ci, va, 2, 4, 1.0, capi.PREC_FP64) as ctx:
        assert ctx.info()["kernel"] == capi.KERNEL_SQA_WAVE
    for n, M in ((1000, 200), (3000, 64)):
        rp, ci, va = gauss_sparse_csr(n, 4, 0x6001)
        with pytest.raises(capi.OjmiError):
            capi.Context.csr(n, rp, ci, va, 2, M, 1.0, capi.PREC_FP64)


def test_sqa_csr_full_anneal_equals_dense_trotter_kernel():
    """A full anneal on EA3D L = 6 (216 spins, +-J), M = 16, 300 sweeps, s
    from 0 to 1: with integer couplings every field sum is exact in both
    orders, so the CSR Trotter kernel and the dense Trotter kernel (fields
    in registers, updated per flip) walk the same trajectory, and both equal
    the oracle's CSR replay."""
    n, rp, ci, va = H.ea3d_csr(6, 0x6006)
    M, R, L, gamma = 16, 32, 300, 1.0
    s = np.linspace(0.0, 1.0, L)
    beta = np.full(L, 5.0)
    mcs = np.ones(L, dtype=np.int64)
    init = np.stack([H.random_pm1(M, n, 120 + r) for r in range(R)])
    E = np.zeros((n + 1, n + 1))
    rows = np.repeat(np.arange(n + 1), np.diff(rp))
    up = rows < ci
    E[rows[up], ci[up]] = va[up]
    out = {}
    for kind, make in (("csr", lambda: capi.Context.csr(n, rp, ci, va, R, M, gamma,
                                                        capi.PREC_FP64)),
                       ("dense", lambda: capi.Context.dense(E, R, M, gamma, capi.PREC_FP64))):
        with make() as ctx:
            assert ctx.info()["kernel"] == (capi.KERNEL_SQA_CSR if kind == "csr"
                                            else capi.KERNEL_SQA_WAVE)
            ctx.set_spins(init)
            ctx.run(beta, mcs, 0xABCD, 0, s=s)
            out[kind] = (ctx.get_spins(), ctx.energies().reshape(R, M))
    np.testing.assert_array_equal(out["csr"][0], out["dense"][0])
    np.testing.assert_array_equal(out["csr"][1], out["dense"][1])
    a, B = orc.expand_transverse(beta, s, mcs, gamma, M)
    want = orc.replay_sqa_csr("fp64", rp, ci, va, n, init, 0xABCD, 0, a, B)
    np.testing.assert_array_equal(out["csr"][0], want)
    best = out["csr"][1].min(axis=1)
    print(f"best energy per spin: min {best.min() / n:.4f} median {np.median(best) / n:.4f}")
    assert best.min() / n < -1.6  # 3D +-J ground states lie near -1.79 per spin
