"""Numerics against scipy (reference test_numerics.py): single-step exponentials of the seed-27
random Hamiltonians in both precisions to eps d^2 (summed absolute error), the debug expm helper,
and a ten-step two-field propagator against the step-by-step product of exponentials."""
import numpy as np
import pytest
import scipy.linalg

import parament_amd as parament

pytestmark = pytest.mark.gpu


def seeded_hamiltonian(dim):
    """H0 = A + A^H with A's real parts U(-1, 1) / dim and imaginary parts U(-1, 1) / 2, drawn from
    numpy's legacy generator at seed 27 (the draw the reference's tests use); H1 = 0."""
    np.random.seed(27)
    a = np.random.uniform(-1, 1, (dim, dim)) / dim + 1j * np.random.uniform(-1, 1, (dim, dim)) / 2
    return a + a.conj().T, np.zeros((dim, dim), dtype=complex)


@pytest.mark.parametrize("precision,dim", [("fp32", 2), ("fp32", 16), ("fp64", 2), ("fp64", 16)])
def test_single_step_matches_scipy(precision, dim):
    h0, h1 = seeded_hamiltonian(dim)
    dt = 0.01
    with parament.Parament(precision) as ctx:
        ctx.set_hamiltonian(h0, h1, use_magnus=False, quadrature_mode="none")
        u = ctx.equiprop(dt, np.zeros(1))
    eps = np.finfo(np.float32 if precision == "fp32" else np.float64).eps
    assert np.sum(np.abs(u - scipy.linalg.expm(-1j * dt * h0))) < eps * dim * dim


@pytest.mark.parametrize("dim", [2, 4])
def test_debug_expm(dim):
    h0, _ = seeded_hamiltonian(dim)
    err = np.sum(np.abs(parament.debug_functions.expm(h0) - scipy.linalg.expm(h0)))
    assert err < np.finfo(np.float32).eps * dim * dim


@pytest.mark.parametrize("dim", [2, 16])
def test_two_fields_ten_steps(dim):
    h0, h1 = seeded_hamiltonian(dim)
    h2 = h0.copy()
    dt = 0.01
    rng = np.random.default_rng(dim)
    c1, c2 = rng.random(10), rng.random(10)
    with parament.Parament("fp32") as ctx:
        ctx.set_hamiltonian(h0, h1, h2, use_magnus=False, quadrature_mode="none")
        u = ctx.equiprop(dt, c1, c2)
    ref = np.eye(dim, dtype=complex)
    for a, b in zip(c1, c2):   # later steps on the left
        ref = scipy.linalg.expm(-1j * dt * (h0 + a * h1 + b * h2)) @ ref
    assert np.linalg.norm(ref - u) < 1e-6
