"""The `cxxjij` binding surface: graphs, schedules, systems, Algorithm run and
get_solution, following the reference's own tests (tests/test_cxx.py,
tests/test_model.py) and binding (openjij/declare.hpp).

CPU tests cover everything that does not anneal; `gpu` tests run the
reference's test_cxx.py flows and replay a cxxjij run on the oracle."""
import json
import os

import numpy as np
import pytest
from scipy import sparse as sp

from oracle import oracle as orc
from tests import helpers as H

cj = pytest.importorskip("openjij_amd").cxxjij

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                  "reference_fixtures.json")))
CASE1 = FIX["testcases"]["1"]  # the 8-spin instance of tests/test_cxx.py:31-67


def fill_case1(g):
    J = np.array(CASE1["J_ext_upper"])
    n = CASE1["n"]
    for i in range(n):
        if J[i, n] != 0:
            g[i] = J[i, n]
        for j in range(i + 1, n):
            if J[i, j] != 0:
                g[i, j] = J[i, j]
    return g


def case1_csr():
    return cj.graph.CSRSparse(sp.csr_matrix(np.array(CASE1["J_ext_upper"])))


# --------------------------------------------------------------------------- graphs
def test_dense_accessors_and_layout():
    g = fill_case1(cj.graph.Dense(8))
    J = np.array(CASE1["J_ext_upper"])
    full = J + np.triu(J, 1).T
    np.testing.assert_array_equal(g.get_interactions(), full)
    assert g[1, 0] == g[0, 1] == J[0, 1]
    assert g[3] == J[3, 8]
    assert g.size() == 8 and g.get_num_spins() == 8


def test_dense_calc_energy_matches_exact_energy():
    g = fill_case1(cj.graph.Dense(8))
    J = np.array(CASE1["J_ext_upper"])
    for seed in range(5):
        s = H.random_pm1(1, 8, seed)[0]
        want = orc.energy_exact(H.symmetric_ext(J), s[None, :])[0]
        assert g.calc_energy(list(s.astype(float))) == pytest.approx(want, abs=1e-12)


def test_dense_set_interaction_matrix_validation():
    g = cj.graph.Dense(3)
    with pytest.raises(RuntimeError):
        g.set_interaction_matrix(np.zeros((3, 4)))
    with pytest.raises(RuntimeError):
        g.set_interaction_matrix(np.zeros((3, 3)))
    m = np.zeros((4, 4))
    m[3, 3] = 1
    m[1, 1] = 0.5
    with pytest.raises(RuntimeError, match="diagonal"):
        g.set_interaction_matrix(m)
    m[1, 1] = 0
    m[3, 3] = 2
    with pytest.raises(RuntimeError, match="unity"):
        g.set_interaction_matrix(m)
    m[3, 3] = 1
    m[0, 2] = -1.5
    m[2, 0] = -1.5
    m[1, 3] = m[3, 1] = 0.25
    g.set_interaction_matrix(m)
    assert g[0, 2] == -1.5 and g[1] == 0.25


def test_sparse_graph_adjacency_and_edge_limit():
    g = cj.graph.Sparse(4, 1)
    assert g.get_num_edges() == 1
    g[0, 1] = -1.0
    g[0, 2] = 2.0  # node 0 is full: stored, but not an edge (sparse.hpp:81-104)
    assert list(g.adj_nodes(0)) == [1]
    assert list(g.adj_nodes(2)) == []
    s = [1, -1, 1, 1]
    assert g.calc_energy(s) == pytest.approx(1.0)  # only J01 counts
    with pytest.raises(IndexError):
        g[1, 3]
    with pytest.raises(IndexError):
        g.adj_nodes(7)


def test_sparse_dense_csr_systems_share_one_matrix():
    d = fill_case1(cj.graph.Dense(8))
    s = fill_case1(cj.graph.Sparse(8))
    c = case1_csr()
    spins = d.gen_spin(3)
    Md = cj.system.make_classical_ising(spins, d).interaction
    Ms = cj.system.make_classical_ising(spins, s).interaction.toarray()
    Mc = cj.system.make_classical_ising(spins, c).interaction.toarray()
    np.testing.assert_array_equal(Md, Ms)
    np.testing.assert_array_equal(Md, Mc)
    for x in range(4):
        v = list(H.random_pm1(1, 8, x)[0].astype(float))
        assert s.calc_energy(v) == pytest.approx(d.calc_energy(v), abs=1e-12)
        assert c.calc_energy(v) == pytest.approx(d.calc_energy(v), abs=1e-12)


def test_csr_reads_upper_triangle_only():
    J = np.array(CASE1["J_ext_upper"])
    full = J + np.triu(J, 1).T
    a = cj.graph.CSRSparse(sp.csr_matrix(full))
    b = cj.graph.CSRSparse(sp.csr_matrix(J))
    np.testing.assert_array_equal(a.get_interactions().toarray(), b.get_interactions().toarray())


@pytest.mark.parametrize("n,seed", [(1, 0), (8, 1234), (100, 7), (1000, 2**32 - 1)])
def test_gen_spin_is_the_reference_stream(n, seed):
    """graph.gen_spin(seed): xorshift128 with the reference's fixed state into
    uniform_int_distribution(0, 1) (include/openjij/graph/graph.hpp:54-65)."""
    g = cj.graph.Dense(n)
    want = orc.ref_gen_spin(n, seed)
    np.testing.assert_array_equal(np.array(g.gen_spin(seed)), want)
    np.testing.assert_array_equal(np.array(g.gen_binary(seed)), (want + 1) // 2)


# --------------------------------------------------------------------------- schedules
def test_schedule_lists_follow_the_builders():
    f = FIX["classical_schedule"]
    sl = cj.utility.make_classical_schedule_list(f["beta_min"], f["beta_max"], f["one_mc_step"],
                                                 f["num_call_updater"])
    betas = np.array([e.updater_parameter.beta for e in sl])
    np.testing.assert_array_equal(
        betas, orc.classical_schedule(f["beta_min"], f["beta_max"], f["num_call_updater"]))
    assert all(e.one_mc_step == f["one_mc_step"] for e in sl)
    t = FIX["transverse_schedule"]
    tl = cj.utility.make_transverse_field_schedule_list(t["beta"], t["one_mc_step"],
                                                        t["num_call_updater"])
    s = np.array([e.updater_parameter.s for e in tl])
    np.testing.assert_array_equal(s, orc.transverse_schedule(t["num_call_updater"]))
    assert s[0] == 0.0 and abs(s[-1] - 1.0) < 1e-12
    assert all(e.updater_parameter.beta == t["beta"] for e in tl)


def test_schedule_objects_and_tuple_lists():
    p = cj.utility.ClassicalUpdaterParameter(2.5)
    cs = cj.utility.ClassicalSchedule((p, 7))
    assert cs.updater_parameter.beta == 2.5 and cs.one_mc_step == 7
    tp = cj.utility.TransverseFieldUpdaterParameter(3.0, 0.25)
    ts = cj.utility.TransverseFieldSchedule((tp, 4))
    assert (ts.updater_parameter.beta, ts.updater_parameter.s, ts.one_mc_step) == (3.0, 0.25, 4)
    lst = cj.utility.make_schedule_list([(0.5, 3), (1.0, 4)])
    assert [(e.updater_parameter.beta, e.one_mc_step) for e in lst] == [(0.5, 3), (1.0, 4)]


# --------------------------------------------------------------------------- systems
def test_classical_system_state_and_errors():
    g = fill_case1(cj.graph.Dense(8))
    s0 = g.gen_spin(1)
    sys = cj.system.make_classical_ising(s0, g)
    assert sys.num_spins == 8
    np.testing.assert_array_equal(sys.spin, list(s0) + [1])
    with pytest.raises(IndexError):
        sys.reset_spins([1] * 7)
    with pytest.raises(ValueError):
        sys.reset_spins([1] * 7 + [0])
    # get_solution multiplies by the dummy spin (get_solution.hpp:39-49)
    sys.spin = np.array(list(s0) + [-1], dtype=float)
    assert cj.result.get_solution(sys) == [-x for x in s0]


def test_transverse_system_shapes_and_errors():
    g = fill_case1(cj.graph.Dense(8))
    with pytest.raises(ValueError):
        cj.system.make_transverse_ising(g.gen_spin(1), g, 1.0, 1)
    with pytest.raises(ValueError):
        cj.system.make_transverse_ising([g.gen_spin(1)], g, 1.0)
    sys = cj.system.make_transverse_ising(g.gen_spin(1), g, 1.5, 4)
    assert sys.trotter_spins.shape == (9, 4)
    assert sys.gamma == 1.5
    sys.gamma = 0.5
    assert sys.gamma == 0.5
    assert np.all(sys.trotter_spins[8] == 1)


def test_transverse_get_solution_is_first_minimum_slice():
    f = FIX["trotter_solution"]
    n, M = f["n"], f["num_trotter"]
    g = cj.graph.Dense(n)
    for i, v in f["h"].items():
        g[int(i)] = v
    for (i, j, v) in f["J"]:
        g[i, j] = v
    rng = np.random.default_rng(0)
    for _ in range(5):
        t = (2 * rng.integers(0, 2, size=(M, n)) - 1).astype(np.int8)
        sys = cj.system.make_transverse_ising([list(map(int, r)) for r in t], g, 1.0)
        want = orc.ref_solution_transverse(np.array(g.get_interactions()), t)
        assert cj.result.get_solution(sys) == list(map(int, want))


# --------------------------------------------------------------------------- GPU flows
def _graphs():
    return {"Dense": fill_case1(cj.graph.Dense(8)), "Sparse": fill_case1(cj.graph.Sparse(8)),
            "CSRSparse": case1_csr()}


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["Dense", "Sparse", "CSRSparse"])
def test_reference_flow_classical(kind):
    """tests/test_cxx.py:224-275: make_classical_schedule_list(0.1, 100, 100, 100)."""
    g = _graphs()[kind]
    sys = cj.system.make_classical_ising(fill_case1(cj.graph.Dense(8)).gen_spin(1234), g)
    assert type(sys).__name__ == "ClassicalIsing_" + kind
    sl = cj.utility.make_classical_schedule_list(0.1, 100.0, 100, 100)
    cj.algorithm.Algorithm_SingleSpinFlip_run(sys, 5678, sl)
    assert cj.result.get_solution(sys) == CASE1["ground_state"]


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["Dense", "Sparse", "CSRSparse"])
def test_reference_flow_transverse(kind):
    """tests/test_cxx.py:419-469: gamma 1, 10 slices, make_transverse_field_schedule_list(10, 100, 100)."""
    g = _graphs()[kind]
    sys = cj.system.make_transverse_ising(fill_case1(cj.graph.Dense(8)).gen_spin(1234), g,
                                          1.0, 10)
    assert type(sys).__name__ == "TransverseIsing_" + kind
    sl = cj.utility.make_transverse_field_schedule_list(10, 100, 100)
    cj.algorithm.Algorithm_SingleSpinFlip_run(sys, 5678, sl)
    assert cj.result.get_solution(sys) == CASE1["ground_state"]


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["Dense", "CSRSparse"])
def test_cxxjij_run_replays_on_the_oracle(kind):
    """The binding hands schedule, seed and spins to the engine unchanged: the
    spins after a cxxjij run equal the FP64 replay of the same seed."""
    g = _graphs()[kind]
    s0 = g.gen_spin(99)
    sys = cj.system.make_classical_ising(s0, g)
    sl = cj.utility.make_classical_schedule_list(0.2, 5.0, 3, 40)
    cj.algorithm.Algorithm_SingleSpinFlip_run(sys, 4242, sl)
    beta = np.array([e.updater_parameter.beta for e in sl])
    mcs = np.array([e.one_mc_step for e in sl], np.int64)
    J = H.symmetric_ext(np.array(CASE1["J_ext_upper"]))
    want = orc.replay_dense("fp64", J, np.array(s0, np.int8)[None, :], 4242, 0,
                            orc.expand_classical(beta, mcs))
    np.testing.assert_array_equal(np.array(sys.spin[:8], np.int8), want[0])


@pytest.mark.gpu
def test_callback_run_with_real_couplings_equals_plain_run():
    """Gaussian couplings (FP64 path): one sweep per device call with a
    callback keeps the device's fields between sweeps, so the trajectory is
    the plain run's, and both equal the replay of the oracle."""
    n = 150
    E = H.sk_gauss(n, 0xCB, field=True)
    g = cj.graph.Dense(n)
    for i in range(n):
        g[i] = E[i, n]
        for j in range(i + 1, n):
            g[i, j] = E[i, j]
    s0 = g.gen_spin(3)
    tl = [(0.2, 4), (1.0, 3), (4.0, 5)]
    a = cj.system.make_classical_ising(s0, g)
    cj.algorithm.Algorithm_SingleSpinFlip_run(a, 77, tl)
    b = cj.system.make_classical_ising(s0, g)
    calls = []
    cj.algorithm.Algorithm_SingleSpinFlip_run(b, 77, tl, lambda sys, beta: calls.append(beta))
    assert len(calls) == 12
    assert list(a.spin) == list(b.spin)
    init = np.array(s0[:n], dtype=np.int8).reshape(1, n)
    beta = np.array([0.2, 1.0, 4.0])
    mcs = np.array([4, 3, 5], np.int64)
    want = orc.replay_dense("fp64", H.symmetric_ext(E), init, 77, 0,
                            orc.expand_classical(beta, mcs))
    assert list(a.spin[:n]) == list(want[0].astype(float))


@pytest.mark.gpu
def test_sparse_system_equals_csr_system():
    s0 = fill_case1(cj.graph.Dense(8)).gen_spin(5)
    out = []
    for g in (fill_case1(cj.graph.Sparse(8)), case1_csr()):
        sys = cj.system.make_classical_ising(s0, g)
        cj.algorithm.Algorithm_SingleSpinFlip_run(sys, 11, [(0.3, 5), (1.0, 5), (3.0, 5)])
        out.append(list(sys.spin))
    assert out[0] == out[1]


@pytest.mark.gpu
def test_run_overloads_callback_and_determinism():
    g = fill_case1(cj.graph.Dense(8))
    s0 = g.gen_spin(7)
    tl = [(0.5, 3), (2.0, 4)]
    seen = []
    a = cj.system.make_classical_ising(s0, g)
    cj.algorithm.Algorithm_SingleSpinFlip_run(
        a, 31, tl, lambda sys, beta: seen.append((beta, list(sys.spin))))
    assert [b for b, _ in seen] == [0.5] * 3 + [2.0] * 4  # one call per sweep (algorithm.hpp:45-53)
    b = cj.system.make_classical_ising(s0, g)
    cj.algorithm.Algorithm_SingleSpinFlip_run(b, 31, tl)
    assert list(a.spin) == list(b.spin) == seen[-1][1]
    c = cj.system.make_classical_ising(s0, g)
    cj.algorithm.Algorithm_SingleSpinFlip_run(c, 31, cj.utility.make_schedule_list(tl))
    assert list(c.spin) == list(b.spin)
    d = cj.system.make_classical_ising(s0, g)
    cj.algorithm.Algorithm_SingleSpinFlip_run(d, tl)  # unseeded overload runs
    assert sorted(set(d.spin)) in ([-1.0, 1.0], [1.0])
    # transverse tuple list with callback: ((beta, s), mcs)
    t = cj.system.make_transverse_ising(s0, g, 1.0, 4)
    params = []
    cj.algorithm.Algorithm_SingleSpinFlip_run(t, 3, [((5.0, 0.2), 2), ((5.0, 0.9), 1)],
                                              lambda sys, p: params.append(tuple(p)))
    assert params == [(5.0, 0.2), (5.0, 0.2), (5.0, 0.9)]
