"""ojmi_energies_static (allocation-free energies) through the C ABI: two calls
in a row on one GPU, the second with a smaller input, each equal bit for bit
to ojmi_energies on the same context; contexts too large for the static
scratch are rejected with OJMI_ERANGE before any device work."""
import ctypes as C

import numpy as np
import pytest

from openjij_amd import capi
from oracle import oracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu


def static_energies(ctx):
    out = np.full((ctx.R, ctx.M) if ctx.M > 1 else ctx.R, np.nan)
    rc = capi.lib().ojmi_energies_static(ctx.h, out.ctypes.data_as(C.c_void_p))
    assert rc == capi.OK, capi.lib().ojmi_last_error().decode()
    return out


def test_static_energies_equal_allocating_entry_point_on_two_calls():
    cases = [  # large first, then smaller: the second call must not see leftovers
        (H.sk_gauss(4096, 0xE005), 2048, 1, capi.PREC_FP32),
        (H.int_dense(300, 12), 5, 1, capi.PREC_INT32),
        (H.sk_gauss(90, 7, field=True), 3, 4, capi.PREC_FP64),
    ]
    for E, R, M, prec in cases:
        n = E.shape[0] - 1
        init = H.random_pm1(R * M, n, R + n).reshape((R, M, n) if M > 1 else (R, n))
        with capi.Context.dense(E, R, M, 1.0, prec) as ctx:
            ctx.set_spins(init)
            got = static_energies(ctx)
            want = ctx.energies()
            again = static_energies(ctx)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(again, want)
        if n <= 300:
            ex = orc.energy_exact(E, init.reshape(R * M, n)).reshape(want.shape)
            assert np.all(np.abs(got - ex) <= 1e-12 * (1 + np.abs(ex)))


def test_static_energies_of_sparse_contexts_on_two_calls():
    """Sparse contexts on the CSR kernel: the static entry point runs the
    level-order energy kernel into the static buffer."""
    for L, R in ((8, 300), (4, 9)):  # larger first
        n, rp, ci, v = H.ea3d_csr(L, 0xA110 + L)
        init = H.random_pm1(R, n, L)
        with capi.Context.csr(n, rp, ci, v, R, precision=capi.PREC_INT32) as ctx:
            assert ctx.info()["kernel"] == capi.KERNEL_CSR_LEVEL
            ctx.set_spins(init)
            got = static_energies(ctx)
            want = ctx.energies()
            again = static_energies(ctx)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(again, want)
        for r in range(R):
            assert got[r] == orc.energy_exact_csr(rp, ci, v, n, init[r])


def test_static_energies_of_sparse_trotter_contexts_on_two_calls():
    """Trotter contexts on the CSR Trotter kernel: one energy per (replica,
    slice) from the same level-order energy kernel."""
    for L, R, M in ((8, 40, 6), (4, 3, 3)):  # larger first
        n, rp, ci, v = H.ea3d_csr(L, 0xA210 + L)
        init = np.stack([H.random_pm1(M, n, L + r) for r in range(R)])
        with capi.Context.csr(n, rp, ci, v, R, M, 1.0, capi.PREC_FP64) as ctx:
            assert ctx.info()["kernel"] == capi.KERNEL_SQA_CSR
            ctx.set_spins(init)
            got = static_energies(ctx)
            want = ctx.energies()
            again = static_energies(ctx)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(again, want)
        for r in range(R):
            for t in range(M):
                assert got[r, t] == orc.energy_exact_csr(rp, ci, v, n, init[r, t])


def test_static_energies_reject_contexts_beyond_the_scratch():
    E = H.sk_pm1(4096, 1)
    with capi.Context.dense(E, 4096, 1, 0.0, capi.PREC_FP64) as ctx:
        out = np.zeros(4096)
        rc = capi.lib().ojmi_energies_static(ctx.h, out.ctypes.data_as(C.c_void_p))
        assert rc == capi.ERANGE
        assert "static buffer" in capi.lib().ojmi_last_error().decode()
        assert not out.any()
