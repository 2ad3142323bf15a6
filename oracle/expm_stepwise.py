"""Stepwise scipy-expm propagator -- TEST INFRASTRUCTURE ONLY (the "truth" oracle).

U = expm(-i tau X_{N-1}) ... expm(-i tau X_0) in complex128, with the same
quadrature-averaged per-slice Hamiltonians X_j that the reference builds
(parament.cpp:486-563).  This is the comparison the reference's own tests make
(test_numerics.py:96, 166-171).  Only tests/, smoke() and bench.py's
cpu_baseline leg may use it.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg

from .equiprop_oracle import slice_hamiltonians


def expm_stepwise(h0, hks, carr, dt: float, mode="none", use_magnus: bool = False,
                  xs=None, tau=None) -> np.ndarray:
    if xs is None:
        xs, tau = slice_hamiltonians(h0, hks, carr, dt, mode, use_magnus)
    d = xs.shape[-1]
    u = np.eye(d, dtype=np.complex128)
    for x in xs:
        u = scipy.linalg.expm(-1j * tau * x) @ u
    return u
