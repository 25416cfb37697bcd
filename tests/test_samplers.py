"""BinaryQuadraticModel, Response and the SASampler / SQASampler front ends,
following the reference's tests/test_model.py and tests/test_sampler.py.

CPU tests cover the model, its conversions and the argument handling; `gpu`
tests sample (every sample call anneals on the GPU)."""
import json
import os

import numpy as np
import pytest

oj = pytest.importorskip("openjij_amd")
cj = oj.cxxjij

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                  "reference_fixtures.json")))

# tests/test_model.py:33-40
H_M = {0: 1, 1: -2}
J_M = {(0, 1): -1, (1, 2): -3, (2, 3): 0.5}
SPINS_M = {0: 1, 1: -1, 2: 1, 3: 1}
Q_M = {(0, 0): 1, (1, 2): -1, (2, 0): -0.2, (1, 3): 3}
BIN_M = {0: 0, 1: 1, 2: 1, 3: 0}


def ising_energy(h, J, s):
    return sum(v * s[i] * s[j] for (i, j), v in J.items()) + sum(v * s[i] for i, v in h.items())


def qubo_energy(Q, x):
    return sum(v * x[i] * x[j] for (i, j), v in Q.items())


def brute_min(bqm):
    """Exhaustive minimum of a small model in its own domain."""
    n = bqm.num_variables
    vals = (-1, 1) if bqm.vartype == oj.SPIN else (0, 1)
    best = None
    for k in range(1 << n):
        x = [vals[(k >> i) & 1] for i in range(n)]
        e = bqm.energy(x)
        best = e if best is None or e < best else best
    return best


# --------------------------------------------------------------------------- model (CPU)
def test_variable_type():
    assert oj.cast_vartype("SPIN") == oj.SPIN
    assert oj.cast_vartype("BINARY") == oj.BINARY


def test_bqm_constructor_and_graph_types():
    bqm = oj.BinaryQuadraticModel(H_M, J_M, "SPIN", sparse=False)
    assert isinstance(bqm.interaction_matrix(), np.ndarray)
    assert bqm.vartype == oj.SPIN
    g, off = bqm.get_cxxjij_ising_graph()
    assert isinstance(g, cj.graph.Dense) and off == 0
    assert oj.BinaryQuadraticModel.from_qubo(Q=Q_M).vartype == oj.BINARY
    gs, _ = oj.BinaryQuadraticModel(H_M, J_M, "SPIN", sparse=True).get_cxxjij_ising_graph()
    assert isinstance(gs, cj.graph.CSRSparse)


def test_interaction_matrix_layout():
    """tests/test_model.py:57-68: upper triangle, h in the last column, corner 1."""
    f = FIX["model_layout"]
    bqm = oj.BinaryQuadraticModel({int(k): v for k, v in f["h"].items()},
                                  {(i, j): v for i, j, v in f["J"]}, "SPIN")
    np.testing.assert_array_equal(bqm.interaction_matrix(), np.array(f["matrix"], float))
    bqm2 = oj.BinaryQuadraticModel({}, {(0, 1): -1.0, (1, 2): -3.0}, "SPIN")
    assert bqm2.get_quadratic() == {(0, 1): -1.0, (1, 2): -3.0}
    np.testing.assert_array_equal(
        bqm2.interaction_matrix(),
        np.array([[0, -1, 0, 0], [0, 0, -3, 0], [0, 0, 0, 0], [0, 0, 0, 1]], float))


@pytest.mark.parametrize("sparse", [False, True])
def test_transfer_to_cxxjij(sparse):
    bqm = oj.BinaryQuadraticModel(H_M, J_M, "SPIN", sparse=sparse)
    g, off = bqm.get_cxxjij_ising_graph()
    I = g.get_interactions()
    I = I.toarray() if hasattr(I, "toarray") else I
    M = bqm.interaction_matrix()
    n = bqm.num_variables
    assert g.size() == n
    for i in range(n + 1):
        for j in range(i + 1, n + 1):
            assert I[i, j] == pytest.approx(M[i, j])
            assert I[i, j] == I[j, i]
    q = oj.BinaryQuadraticModel.from_qubo(Q=Q_M, sparse=sparse)
    gq, offq = q.get_cxxjij_ising_graph()
    sq = q.change_vartype("SPIN", inplace=False)
    Iq = gq.get_interactions()
    Iq = Iq.toarray() if hasattr(Iq, "toarray") else Iq
    Mq = sq.interaction_matrix()
    for i in range(n + 1):
        for j in range(i + 1, n + 1):
            assert Iq[i, j] == pytest.approx(Mq[i, j])
    assert offq == pytest.approx(sq.offset)


def test_bqm_energies_closed_forms():
    bqm = oj.BinaryQuadraticModel(H_M, J_M, "SPIN")
    assert bqm.energy(SPINS_M) == ising_energy(H_M, J_M, SPINS_M)
    q = oj.BinaryQuadraticModel.from_qubo(Q=Q_M)
    assert q.energy(BIN_M) == qubo_energy(Q_M, BIN_M)
    e_bin = q.energy({0: 1, 1: 1, 2: 0, 3: 1})
    q.change_vartype("SPIN")
    assert e_bin == pytest.approx(q.energy({0: 1, 1: 1, 2: -1, 3: 1}))
    # every configuration: BINARY energy == SPIN energy of the mapped state
    q2 = oj.BinaryQuadraticModel.from_qubo(Q=Q_M)
    s2 = q2.change_vartype("SPIN", inplace=False)
    for k in range(16):
        x = [(k >> i) & 1 for i in range(4)]
        assert q2.energy(x) == pytest.approx(s2.energy([2 * v - 1 for v in x]))


def test_energy_consistency_dense_sparse_graphs():
    spins = {0: -1, 1: -1, 2: -1, 3: -1}
    for sparse in (False, True):
        bqm = oj.BinaryQuadraticModel(H_M, J_M, vartype="SPIN", sparse=sparse)
        g, _ = bqm.get_cxxjij_ising_graph()
        assert g.calc_energy([spins[i] for i in range(4)]) == pytest.approx(bqm.energy(spins))


def test_string_labels_and_numpy_qubo():
    bqm = oj.BinaryQuadraticModel.from_ising({"b": 1.0}, {("a", "c"): -2.0})
    assert bqm.variables == ["a", "b", "c"]
    Q = np.array([[-1, -1, 0], [0, -1, 0], [0, 0, 1]], float)
    q = oj.BinaryQuadraticModel.from_numpy_matrix(Q)
    assert q.energy([1, 1, 0]) == pytest.approx(-3.0)


# --------------------------------------------------------------------------- samplers (CPU)
def test_constructor_keywords_are_rejected():
    """tests/test_sampler.py:263-279."""
    for kw in ("num_reads", "num_sweeps", "beta_min", "beta_max"):
        with pytest.raises(TypeError):
            oj.SASampler(**{kw: 10})
    for kw in ("num_reads", "num_sweeps", "beta", "trotter"):
        with pytest.raises(TypeError):
            oj.SQASampler(**{kw: 10})


def test_invalid_arguments_raise_before_any_device_work():
    sa = oj.SASampler()
    with pytest.raises(ValueError):
        sa.sample_ising({0: 1}, {}, updater="swendsen wang")
    with pytest.raises(ValueError):
        sa.sample_ising({0: 1}, {(0, 1): 1}, initial_state=[1, 1, 1])
    with pytest.raises(RuntimeError):
        sa.sample_ising({0: 1}, {(0, 1): 1}, initial_state=[1, 0])
    sqa = oj.SQASampler()
    with pytest.raises(ValueError):
        sqa.sample_ising({0: 1}, {(0, 1): 1}, schedule=[[1.5, 10]])
    with pytest.raises(ValueError):
        sqa.sample_ising({0: 1}, {(0, 1): 1}, trotter=1)


# --------------------------------------------------------------------------- samplers (GPU)
def _check_basic(res, n_reads=1):
    assert len(res.states) == n_reads and len(res.energies) == n_reads
    assert len(res.record.sample) == n_reads and len(res.record.energy) == n_reads
    assert all(isinstance(float(e), float) for e in res.energies)


def _run_fixture_calls(sampler, init_state=None, init_q_state=None, schedule=None):
    """tests/test_sampler.py:52-63, plus the ground states the fixtures hold."""
    fi, fq = FIX["sampler_ising"], FIX["sampler_qubo"]
    h = {int(k): v for k, v in fi["h"].items()}
    J = {(i, j): v for i, j, v in fi["J"]}
    res = sampler.sample_ising(h, J, schedule=schedule, initial_state=init_state, seed=1,
                               num_reads=8)
    _check_basic(res, 8)
    assert res.first.energy == pytest.approx(fi["ground_energy"])
    Q = {(i, j): v for i, j, v in fq["Q"]}
    res = sampler.sample_qubo(Q, initial_state=init_q_state, schedule=schedule, seed=2,
                              num_reads=8)
    _check_basic(res, 8)
    assert res.first.energy == pytest.approx(fq["ground_energy"])
    assert list(res.states[int(np.argmin(res.energies))]) == fq["ground_state"]
    Qm = np.zeros((5, 5))
    for (i, j), v in Q.items():
        Qm[i, j] += v
    res = sampler.sample_qubo(Qm, initial_state=init_q_state, schedule=schedule, seed=2)
    _check_basic(res, 1)


@pytest.mark.gpu
def test_sa_reference_calls():
    sa = oj.SASampler()
    _run_fixture_calls(sa)
    _run_fixture_calls(sa, init_state=[1] * 5, init_q_state=[1] * 5)
    _run_fixture_calls(sa, init_state={i: 1 for i in range(5)})
    _run_fixture_calls(sa, init_state={i: 1 for i in range(5)},
                       schedule=[[0.1, 10], [1, 10], [10, 10]])
    _run_fixture_calls(sa, init_state={i: 1 for i in range(5)},
                       schedule=[(0.1, 10), (1, 10), (10, 10)])


@pytest.mark.gpu
@pytest.mark.parametrize("sparse", [False, True])
@pytest.mark.parametrize("precision", [None, "fp32", "fp64"])
def test_sa_antiferromagnet_ground_state(sparse, precision):
    """tests/test_sampler.py:41-47: the 30-spin chain with h_0 = -10."""
    f = FIX["sampler_afm"]
    h = {int(k): v for k, v in f["h"].items()}
    J = {(i, i + 1): f["J"] for i in range(f["N"] - 1)}
    res = oj.SASampler().sample_ising(h, J, sparse=sparse, seed=1, num_reads=100,
                                      precision=precision)
    _check_basic(res, 100)
    assert res.first.sample == {i: s for i, s in enumerate(f["ground_state"])}
    bqm = oj.BinaryQuadraticModel.from_ising(h, J)
    np.testing.assert_allclose(res.energies, bqm.energies(res.states), rtol=0, atol=1e-9)


@pytest.mark.gpu
def test_sa_seed_reproducible_and_reads_distinct():
    f = FIX["sampler_afm"]
    h = {int(k): v for k, v in f["h"].items()}
    J = {(i, i + 1): f["J"] for i in range(f["N"] - 1)}
    a = oj.SASampler().sample_ising(h, J, seed=5, num_reads=16, num_sweeps=5)
    b = oj.SASampler().sample_ising(h, J, seed=5, num_reads=16, num_sweeps=5)
    np.testing.assert_array_equal(a.states, b.states)
    assert len({tuple(s) for s in a.states}) > 1  # short anneal: reads differ


@pytest.mark.gpu
def test_sa_reinitialize_false_continues_from_previous_read():
    h, J = {0: -1.0}, {(0, 1): -1.0, (1, 2): -1.0}
    res = oj.SASampler().sample_ising(h, J, num_reads=4, reinitialize_state=False, seed=3,
                                      initial_state=[1, 1, 1], num_sweeps=10)
    _check_basic(res, 4)
    assert all(list(s) == [1, 1, 1] for s in res.states)


@pytest.mark.gpu
def test_sqa_reference_calls():
    sqa = oj.SQASampler()
    _run_fixture_calls(sqa)
    _run_fixture_calls(sqa, init_state=[1] * 5, init_q_state=[1] * 5)
    _run_fixture_calls(sqa, init_state={i: 1 for i in range(5)},
                       schedule=[[0.1, 10], [0.5, 10], [0.9, 10]])
    _run_fixture_calls(sqa, init_state={i: 1 for i in range(5)},
                       schedule=[(0.1, 0.1, 10), (0.5, 1, 10), (0.9, 10, 10)])
    res = sqa.sample_ising({0: -1.0}, {(0, 1): -1.0}, num_reads=3, trotter=6, seed=4)
    tr = res.info["system"]
    assert len(tr) == 3 and np.asarray(tr[0]["trotter_state"]).shape == (6, 2)
    assert len(tr[0]["trotter_energies"]) == 6


@pytest.mark.gpu
def test_sqa_antiferromagnet_ground_state():
    f = FIX["sampler_afm"]
    h = {int(k): v for k, v in f["h"].items()}
    J = {(i, i + 1): f["J"] for i in range(f["N"] - 1)}
    res = oj.SQASampler().sample_ising(h, J, seed=1, num_reads=100)
    _check_basic(res, 100)
    assert res.first.sample == {i: s for i, s in enumerate(f["ground_state"])}


@pytest.mark.gpu
def test_negative_interactions_and_two_variable_cases():
    """tests/test_sampler.py:179-183, 231-236, 281-329."""
    for s in (oj.SASampler(), oj.SQASampler()):
        s.sample_ising({}, {(0, 1): -1})
        s.sample_ising({2: -1}, {(0, 1): -1})
    sa = oj.SASampler()
    for small in (0.0, 1e-7):
        q1 = {(0, 0): 1, (1, 1): 1}
        q2 = {(0, 0): -1, (1, 1): -1}
        if small:
            q1[(0, 1)] = small
            q2[(0, 1)] = small
        assert list(sa.sample_qubo(q1, seed=1).states[0]) == [0, 0]
        assert list(sa.sample_qubo(q2, seed=1).states[0]) == [1, 1]
    sa.sample_qubo({}, seed=1)
    sa.sample_qubo({(0, 0): 0, (1, 1): 0}, seed=1)


@pytest.mark.gpu
def test_empty_models():
    """tests/test_sampler.py:251-261."""
    for s in (oj.SASampler(), oj.SQASampler()):
        for sparse in (True, False):
            assert len(s.sample_ising({}, {}, sparse=sparse).first.sample) == 0
            assert len(s.sample_qubo(Q={}, sparse=sparse).first.sample) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("n,seed", [(10, 1), (14, 2)])
def test_sa_random_models_reach_brute_force_minimum(n, seed):
    rng = np.random.default_rng(seed)
    h = {i: float(rng.normal()) for i in range(n)}
    J = {(i, j): float(rng.normal()) for i in range(n) for j in range(i + 1, n)
         if rng.random() < 0.6}
    bqm = oj.BinaryQuadraticModel.from_ising(h, J)
    res = oj.SASampler().sample_ising(h, J, num_reads=32, num_sweeps=2000, seed=seed)
    assert res.first.energy == pytest.approx(brute_min(bqm), abs=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [None, "fp64"])
def test_sample_batch_reaches_every_minimum(precision):
    rng = np.random.default_rng(12)
    models = []
    for n in (3, 9, 14, 6, 12):
        Q = {(i, j): int(rng.integers(-8, 9)) for i in range(n) for j in range(i, n)
             if rng.random() < 0.5}
        models.append(oj.BinaryQuadraticModel.from_qubo(Q))
    models.insert(2, oj.BinaryQuadraticModel.from_qubo({}))
    models.append(oj.BinaryQuadraticModel.from_ising({0: 0.5}, {(0, 1): -1.25, (1, 2): 0.75}))
    res = oj.SASampler().sample_batch(models, num_reads=16, num_sweeps=500, seed=3,
                                      precision=precision)
    assert len(res) == len(models)
    for bqm, r in zip(models, res):
        assert len(r.states) == 16
        if bqm.num_variables == 0:
            assert r.states.shape == (16, 0)
            continue
        np.testing.assert_allclose(r.energies, bqm.energies(r.states), atol=1e-9)
        assert r.first.energy == pytest.approx(brute_min(bqm), abs=1e-9)
