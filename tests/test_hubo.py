"""SASampler.sample_hubo for polynomial models that reduce to degree <= 2
(openjij/sampler/base_sa_sample_hubo.py:30-140,
include/openjij/sampler/sa_sampler.hpp:27-293): model reduction, energy
difference bounds, beta lists and argument errors on the CPU; on the GPU, the
same trajectories as the oracle replay with the chosen rule, and the ground
states of the reference's test models."""
import itertools
import math

import numpy as np
import pytest

import openjij_amd as oj
from openjij_amd.sampler.hubo import QuadraticPolynomialModel, generate_beta_list
from oracle import oracle as orc
from tests import helpers as H


def brute_energies(model):
    n = model.n
    vals = [-1, 1] if model.vartype == oj.SPIN else [0, 1]
    states = np.array(list(itertools.product(vals, repeat=n)))
    return states, model.energies(states)


def direct_energy(J, vartype, sample):
    """sum over the input dict of value * product of the sample's values."""
    e = 0.0
    for key, v in J.items():
        p = 1
        for x in key:
            p *= sample.get(x, 1)  # a SPIN variable only in s^2 terms: any value
        e += v * p
    return e


@pytest.mark.parametrize("vartype", ["SPIN", "BINARY"])
def test_model_energies_equal_the_input_polynomial(vartype):
    J = {(0,): -1.0, (0, 1): -1.0, (1, 1): 3.0, (2,): 0.5, (): 2.0, ("a", "a"): 0.25,
         (2, 0): 0.75, (1, 2, 2): 1.5}
    if vartype == "BINARY":
        J.pop(("a", "a"))
        J[("a",)] = 0.25
    m = QuadraticPolynomialModel(J, vartype)
    states, en = brute_energies(m)
    for s, e in zip(states, en):
        sample = dict(zip(m.variables, s))
        assert e == pytest.approx(direct_energy(J, vartype, sample), abs=1e-12)
    # the engine's Ising layout has the same energies up to the dropped constant
    ext, const = m.ising_ext()
    n = m.n
    for s, e in zip(states, en):
        sp = s if vartype == "SPIN" else 2 * s - 1
        e_is = (sum(ext[i, j] * sp[i] * sp[j] for i in range(n) for j in range(i + 1, n))
                + sum(ext[i, n] * sp[i] for i in range(n)) + const)
        assert e_is == pytest.approx(e, abs=1e-12)


def test_energy_difference_bounds_follow_the_reference_rule():
    """ising_polynomial_model.hpp:119-145: min over terms of 2|v|, max over
    variables of 2 sum |v| of the terms containing it; binary without the 2;
    terms below 1e-8 * max |v| are ignored."""
    J = {(0,): 1.0, (0, 1): -3.0, (1, 2): 0.5, (2,): 1e-12 * 0 + 2e-9}
    m = QuadraticPolynomialModel(J, "SPIN")
    # (2,) = 2e-9 < 1e-8 * 3 is ignored
    assert m.energy_difference_bounds() == (1.0, 8.0)
    mb = QuadraticPolynomialModel(J, "BINARY")
    assert mb.energy_difference_bounds() == (0.5, 4.0)
    assert QuadraticPolynomialModel({(): 1.0}, "SPIN").energy_difference_bounds() == (0.0, 0.0)


def test_beta_lists_closed_forms():
    g = generate_beta_list("GEOMETRIC", 0.1, 10.0, 5)
    np.testing.assert_allclose(g, 0.1 * 10.0 ** (np.arange(5) / 2), rtol=1e-14)
    lin = generate_beta_list("LINEAR", 0.5, 2.5, 5)
    np.testing.assert_array_equal(lin, [0.5, 1.0, 1.5, 2.0, 2.5])
    np.testing.assert_array_equal(generate_beta_list("LINEAR", 3.0, 9.0, 1), [3.0])
    # the same repeated multiplication the oracle's schedule builder uses
    np.testing.assert_array_equal(g, orc.classical_schedule(0.1, 10.0, 5))


def test_sample_hubo_argument_errors():
    s = oj.SASampler()
    J = {(0, 1): -1.0}
    with pytest.raises(ValueError):
        s.sample_hubo(J, "INTEGER")
    for kw, msg in ((dict(updater="SUWA_TODO"), "Unknown UpdateMethod"),
                    (dict(updater="FOO"), "Invalid update_method"),
                    (dict(random_number_engine="PCG"), "Invalid random_number_engine"),
                    (dict(temperature_schedule="EXP"), "Invalid temperature_schedule"),
                    (dict(num_sweeps=0), "num_sweeps"), (dict(num_reads=0), "num_reads"),
                    (dict(beta_min=-1.0), "beta_min")):
        with pytest.raises(RuntimeError, match=msg):
            s.sample_hubo(J, "SPIN", **kw)
    with pytest.raises(NotImplementedError):
        s.sample_hubo({(0, 1, 2): 1.0}, "SPIN")
    with pytest.raises(NotImplementedError):
        s.sample_hubo(J, "SPIN", updater="k-local")


# --- GPU -----------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("updater", ["METROPOLIS", "HEAT_BATH"])
@pytest.mark.parametrize("schedule", ["GEOMETRIC", "LINEAR"])
def test_sample_hubo_matches_the_oracle_replay(updater, schedule):
    """The GPU run of sample_hubo equals the oracle replay of the same Ising
    layout, initial states, beta list, seed and rule, bit for bit."""
    from openjij_amd.sampler.hubo import UPDATE_METHODS
    from openjij_amd.sampler.sampler import initial_states

    E = H.sk_gauss(40, 0x70B0, field=True)
    J = {}
    for i in range(40):
        J[(i,)] = E[i, 40]
        for j in range(i + 1, 40):
            J[(i, j)] = E[i, j]
    R, L, seed = 6, 60, 4321
    res = oj.SASampler().sample_hubo(J, "SPIN", num_sweeps=L, num_reads=R, updater=updater,
                                     temperature_schedule=schedule, seed=seed,
                                     precision="fp64")
    m = QuadraticPolynomialModel(J, "SPIN")
    ext, _ = m.ising_ext()
    info = res.info["schedule"]
    betas = generate_beta_list(schedule, info["beta_min"], info["beta_max"], L)
    init = initial_states(R, 40, seed, 0x5D)
    want = orc.replay_dense("fp64", H.symmetric_ext(ext), init, seed, 0, -1.0 / betas,
                            rule=UPDATE_METHODS[updater])
    np.testing.assert_array_equal(res.states, want)
    np.testing.assert_allclose(res.energies, m.energies(want), rtol=0, atol=1e-12)
    dmin, dmax = m.energy_difference_bounds()
    assert info["beta_min"] == math.log(2.0) / dmax
    assert info["beta_max"] == math.log(100.0) / dmin


@pytest.mark.gpu
@pytest.mark.parametrize("updater", ["METROPOLIS", "HEAT_BATH"])
@pytest.mark.parametrize("vartype", ["SPIN", "BINARY"])
def test_sample_hubo_finds_brute_force_ground_states(updater, vartype):
    rng = np.random.default_rng(11)
    n = 14
    J = {(i,): float(rng.integers(-4, 5)) for i in range(n)}
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < 0.5:
                J[(i, j)] = float(rng.integers(-4, 5))
    m = QuadraticPolynomialModel(J, vartype)
    _, en = brute_energies(m)
    res = oj.SASampler().sample_hubo(J, vartype, num_sweeps=400, num_reads=32,
                                     updater=updater, seed=5)
    assert res.first.energy == pytest.approx(en.min(), abs=1e-9)
    assert set(np.unique(res.states)) <= ({-1, 1} if vartype == "SPIN" else {0, 1})
    for s, e in zip(res.states, res.energies):
        assert e == pytest.approx(direct_energy(J, vartype, dict(zip(m.variables, s))),
                                  abs=1e-9)
    assert res.info["schedule"]["update_method"] == updater
    assert res.info["execution_time"] > 0


@pytest.mark.gpu
def test_sample_hubo_reference_docstring_models():
    """The examples of sample_hubo's docstring with the degree-3 term
    replaced by degree <= 2 ones (openjij/sampler/sa_sampler.py:519-527)."""
    s = oj.SASampler()
    r = s.sample_hubo({(0,): -1, (0, 1): -1, (1, 2): 1}, "SPIN", num_reads=8, seed=1)
    assert r.first.energy == pytest.approx(-3.0)
    r = s.sample_hubo({(0,): -1, (0, 1): -1, (1, 2): 1}, "BINARY", num_reads=8, seed=1)
    assert r.first.energy == pytest.approx(-2.0)
