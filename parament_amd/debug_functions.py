"""Debug helpers (reference: src/python/pyparament/parament/debug_functions.py:22-38)."""
import numpy as np

from . import paramentlib
from . import parament


def expm(m):
    """Compute the exponential of the matrix `m` via one equiprop slice.

    H0 = 1j*m and H1 = m with a zero control give exp(-i * 1j*m) = exp(m).
    Only works for matrices with small norms.
    """
    dt = 1.0
    with parament.Parament() as gpu_runner:
        gpu_runner.set_hamiltonian(1j * m, m, use_magnus=False, quadrature_mode='none')
        expm_gpu = gpu_runner.equiprop(dt, np.zeros(1))
    return expm_gpu


def device_info():
    """Print information about the available devices to stdout."""
    paramentlib.lib.device_info()
