"""Wrapper life cycle (reference test_wrapper.py)."""
import numpy as np
import pytest

import parament_amd as parament

H0 = np.array([[1, 0], [0, -1]])
H1 = np.array([[0, 1], [1, 0]])


@pytest.mark.gpu
def test_set_hamiltonian_and_destroy():
    ctx = parament.Parament()
    ctx.set_hamiltonian(H0, H1)
    ctx.destroy()


@pytest.mark.gpu
def test_context_manager_destroys():
    with parament.Parament() as ctx:
        ctx.set_hamiltonian(H0, H1)
        ctx.equiprop(1.0, np.zeros(1))
    assert ctx._handle is None


@pytest.mark.gpu
def test_device_info():
    parament.device_info()
