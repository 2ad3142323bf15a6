"""Error behaviour of the Python class (reference test_error.py: the same messages and exception
types)."""
import numpy as np
import pytest

import parament_amd as parament


def test_use_after_destruction():
    ctx = parament.Parament()
    ctx.destroy()
    with pytest.raises(RuntimeError, match="Attempting to use a context that has been destroyed"):
        ctx.set_hamiltonian(np.eye(2), np.eye(2))


def test_propagate_without_hamiltonian():
    with parament.Parament() as ctx:
        with pytest.raises(RuntimeError, match="No hamiltonian set"):
            ctx.equiprop(1.0)


@pytest.mark.parametrize("mode", ["none", "midpoint"])
def test_magnus_needs_simpson(mode):
    with parament.Parament() as ctx:
        with pytest.raises(ValueError, match="Invalid quadrature selection"):
            ctx.set_hamiltonian(np.eye(2), np.eye(2), use_magnus=True, quadrature_mode=mode)


@pytest.mark.gpu
def test_more_amplitudes_than_hamiltonians():
    with parament.Parament() as ctx:
        ctx.set_hamiltonian(np.eye(2), np.eye(2))
        with pytest.raises(ValueError, match="Got 2 amplitude arrays, but there are only 1 Hamiltonians."):
            ctx.equiprop(1.0, np.zeros(4), np.zeros(4))


@pytest.mark.gpu
def test_amplitude_arrays_of_different_length():
    with parament.Parament() as ctx:
        ctx.set_hamiltonian(np.eye(2), np.eye(2), np.eye(2))
        with pytest.raises(ValueError, match="All amplitude arrays must have the same length."):
            ctx.equiprop(1.0, np.zeros(4), np.zeros(5))
