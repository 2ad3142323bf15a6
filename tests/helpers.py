"""Shared input generators for the tests (seeded, documented distributions)."""
import numpy as np


def gue(rng, d, norm=1.0):
    """Hermitian (G + G^H)/2 with complex N(0,1) entries, scaled to spectral norm `norm`."""
    g = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    h = (g + g.conj().T) / 2
    return h * (norm / np.linalg.norm(h, 2))


def random_hamiltonians(dim):
    """The reference's test draw (test_numerics.py:24-29)."""
    np.random.seed(27)
    h0 = np.random.uniform(-1, 1, (dim, dim)) / dim + 1j * np.random.uniform(-1, 1, (dim, dim)) / 2
    h0 = h0 + h0.conj().T
    return h0, np.zeros((dim, dim), dtype=complex)


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def as_c64_inputs(h0, hks, c):
    """Round inputs to what a complex64 caller hands over, returned in complex128."""
    r = lambda x: np.asarray(x).astype(np.complex64).astype(np.complex128)
    return r(h0), [r(h) for h in hks], r(c)
