"""Debug helpers (reference: src/python/pyparament/parament/debug_functions.py:22-38)."""
import numpy as np

from . import paramentlib
from . import parament


def expm(m):
    """exp(m) of a square matrix of small norm, evaluated on the GPU as one equiprop slice.

    One slice of length 1 under the drift 1j*m propagates exp(-1j * 1j*m) = exp(m).  The matrix
    is loaded as a zero-amplitude control as well, so that the norm bound (and with it the
    Chebyshev order) is the one the reference's helper uses.  Complex64 result.
    """
    m = np.asarray(m)
    ctx = parament.Parament('fp32')
    try:
        ctx.set_hamiltonian(1j * m, m, quadrature_mode='none')
        return ctx.equiprop(1.0, np.zeros(1))
    finally:
        ctx.destroy()


def device_info():
    """Print the visible HIP devices (name, architecture, compute units, memory) to stdout."""
    paramentlib.lib.device_info()
