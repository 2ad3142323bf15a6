"""Gate-fidelity gradient oracle -- TEST INFRASTRUCTURE ONLY.

F(c) = |Tr(T^H U(c))|^2 / d^2 with U = expm(-i tau X_{N-1}) ... expm(-i tau X_0) (the exact
exponential, not the Chebyshev approximant), for real control samples c.  dU_j / dc'_kj is
the Frechet derivative L(A_j, -i tau H_k) of expm at A_j = -i tau X_j, taken from the
block-matrix identity expm([[A, E], [0, A]]) = [[e^A, L(A, E)], [0, e^A]] (Van Loan 1978), in
complex128.  The quadrature's chain rule maps dF/dc' to the raw samples: none -> identity,
midpoint -> weights 1/2, 1/2; Simpson -> 1/6, 4/6, 1/6 (control_expansion.cu:105-170).
With Magnus slices (control_expansion.cu:27-60) a slice's exponent depends on its three samples
through the Simpson means, the commutator-with-H0 terms and the pair terms, each affine in any
one sample; dX_j/dc for a sample is therefore X_j(c + e) - X_j(c) exactly, taken from the same
coefficient function the equiprop oracle uses.  ``slices`` restricts the sum to a window of
slice indices (a contiguous block of a long sequence): entries whose samples lie wholly inside
the window are then exact, the others hold only the window's share.
No reference counterpart (the reference has no gradients); pinned in tests/test_oracle.py
against scipy.linalg.expm_frechet and against central finite differences of F.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg

from .equiprop_oracle import (_MODE_BY_NAME, QUAD_MIDPOINT, QUAD_NONE, QUAD_SIMPSON, magnus_coefficients,
                              magnus_operators, quadrature, slice_count)


def frechet_block(a: np.ndarray, e: np.ndarray) -> np.ndarray:
    d = a.shape[0]
    big = np.zeros((2 * d, 2 * d), dtype=np.complex128)
    big[:d, :d] = a
    big[d:, d:] = a
    big[:d, d:] = e
    return scipy.linalg.expm(big)[:d, d:]


def fidelity(h0, hks, c, dt, mode, target, use_magnus=False) -> float:
    from .expm_stepwise import expm_stepwise
    u = expm_stepwise(h0, hks, c, dt, mode, use_magnus)
    g = np.trace(np.asarray(target).conj().T @ u)
    return float(abs(g) ** 2 / target.shape[0] ** 2)


def _magnus_gradient(h0, hks, c, dt, target, slices):
    """(F, dF/dc) for Magnus slices: per slice j and control sample p in {2j, 2j+1, 2j+2}, the
    exponent's derivative is the difference of the slice's Magnus coefficients with that sample
    raised by one (each coefficient is affine in a single sample)."""
    h0 = np.asarray(h0, np.complex128)
    hks = [np.asarray(h, np.complex128) for h in hks]
    c = np.atleast_2d(np.asarray(c, np.float64))
    amps, pts = c.shape
    d = h0.shape[0]
    tau = 2 * dt
    full = np.zeros((len(hks), pts))
    full[:amps] = c
    ops = list(hks) + magnus_operators(h0, hks)
    coef = magnus_coefficients(full, tau)
    n = coef.shape[1]
    xs = [h0 + sum(coef[t, j] * ops[t] for t in range(len(ops))) for j in range(n)]
    us = [scipy.linalg.expm(-1j * tau * x) for x in xs]
    pre = [np.eye(d, dtype=complex)]
    for j in range(n - 1):
        pre.append(us[j] @ pre[-1])
    suf = [None] * n
    suf[n - 1] = np.eye(d, dtype=complex)
    for j in range(n - 2, -1, -1):
        suf[j] = suf[j + 1] @ us[j + 1]
    u = suf[0] @ us[0] @ pre[0]
    th = np.asarray(target, np.complex128).conj().T
    g = np.trace(th @ u)
    dg = np.zeros((amps, pts), dtype=np.complex128)
    for j in (range(n) if slices is None else slices):
        local = full[:, 2 * j:2 * j + 3]
        base = magnus_coefficients(local, tau)[:, 0]
        for k in range(amps):
            for o in range(3):
                bumped = local.copy()
                bumped[k, o] += 1.0
                dcoef = magnus_coefficients(bumped, tau)[:, 0] - base
                dx = sum(dcoef[t] * ops[t] for t in range(len(ops)) if dcoef[t] != 0)
                lk = frechet_block(-1j * tau * xs[j], -1j * tau * dx)
                dg[k, 2 * j + o] += np.trace(th @ suf[j] @ lk @ pre[j])
    return float(abs(g) ** 2 / d ** 2), 2 * np.real(np.conj(g) * dg) / d ** 2


def fidelity_gradient(h0, hks, c, dt, mode, target, use_magnus=False, slices=None):
    """(F, dF/dc) with dF/dc shaped like c (amps x pts); ``slices``: the slice indices whose
    contributions are summed (default: all)."""
    if isinstance(mode, str):
        mode = _MODE_BY_NAME[mode]
    if use_magnus:
        return _magnus_gradient(h0, hks, c, dt, target, slices)
    h0 = np.asarray(h0, np.complex128)
    hks = [np.asarray(h, np.complex128) for h in hks]
    c = np.atleast_2d(np.asarray(c, np.float64))
    amps, pts = c.shape
    d = h0.shape[0]
    tau = 2 * dt if mode == QUAD_SIMPSON else dt
    cq = quadrature(np.vstack([c, np.zeros((len(hks) - amps, pts))]), mode).real
    n = cq.shape[1]
    xs = [h0 + sum(cq[k, j] * hks[k] for k in range(len(hks))) for j in range(n)]
    us = [scipy.linalg.expm(-1j * tau * x) for x in xs]
    pre = [np.eye(d, dtype=complex)]
    for j in range(n - 1):
        pre.append(us[j] @ pre[-1])
    suf = [None] * n
    suf[n - 1] = np.eye(d, dtype=complex)
    for j in range(n - 2, -1, -1):
        suf[j] = suf[j + 1] @ us[j + 1]
    u = suf[0] @ us[0] @ pre[0]
    th = np.asarray(target, np.complex128).conj().T
    g = np.trace(th @ u)
    dgq = np.zeros((amps, n), dtype=np.complex128)
    for j in (range(n) if slices is None else slices):
        a = -1j * tau * xs[j]
        for k in range(amps):
            lk = frechet_block(a, -1j * tau * hks[k])
            dgq[k, j] = np.trace(th @ suf[j] @ lk @ pre[j])
    dfq = 2 * np.real(np.conj(g) * dgq) / d ** 2
    grad = np.zeros((amps, pts))
    if mode == QUAD_NONE:
        grad[:, :n] = dfq
    elif mode == QUAD_MIDPOINT:
        grad[:, :n] += 0.5 * dfq
        grad[:, 1:n + 1] += 0.5 * dfq
    else:
        grad[:, 0:2 * n:2] += dfq / 6
        grad[:, 1:2 * n:2] += 4 * dfq / 6
        grad[:, 2:2 * n + 1:2] += dfq / 6
    return float(abs(g) ** 2 / d ** 2), grad
