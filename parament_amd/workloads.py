"""Seeded synthetic inputs for the reference configurations of the equiprop path.

Each generator returns ``(H0, [H_k], carr, dt, quadrature)`` with complex128 arrays; the
distributions are the ones the benchmark and the tests are specified on:

* ``dim8_gue``          -- dim 8, GUE H0 and one GUE control (complex N(0,1) entries,
                           symmetrised, spectral norm 1), 4096 steps, dt 1e-3, controls
                           U[-1, 1], no quadrature.
* ``heisenberg6``       -- dim 64: open 6-spin Heisenberg chain sum_i J_i (XX + YY + ZZ),
                           J ~ U[0.5, 1.5]; controls sum_i X_i, sum_i Y_i, sum_i Z_i with
                           sum-of-sines waveforms (4 tones, random phases, max |c| = 2);
                           65,536 steps, dt 5e-3, Simpson quadrature.
* ``ising8``            -- dim 256: 8-qubit transverse-field Ising chain sum Z_i Z_{i+1} +
                           sum h_i X_i, h ~ U[0, 2]; 8 local X_i controls with Gaussian-
                           smoothed N(0, 1) waveforms; 16,384 steps, dt 2e-3.
* ``gue32_grape``       -- dim 32, GUE H0 plus 4 GUE controls, 4096 steps, dt 2e-3, controls
                           U[-1, 1], no quadrature, plus a Haar-random target unitary.
"""
from __future__ import annotations

import numpy as np

SX = np.array([[0, 1], [1, 0]], dtype=complex)
SY = np.array([[0, -1j], [1j, 0]], dtype=complex)
SZ = np.array([[1, 0], [0, -1]], dtype=complex)


def gue(rng, d, norm=1.0):
    g = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    h = (g + g.conj().T) / 2
    return h * (norm / np.linalg.norm(h, 2))


def site_op(op, i, n):
    m = np.array([[1.0 + 0j]])
    for k in range(n):
        m = np.kron(m, op if k == i else np.eye(2))
    return m


def haar_unitary(rng, d):
    z = (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return q * ph[None, :]


def sum_of_sines(rng, amps, pts, dt, tones=4, amax=2.0):
    t = np.arange(pts) * dt
    c = np.zeros((amps, pts))
    for k in range(amps):
        for _ in range(tones):
            c[k] += np.sin(2 * np.pi * rng.uniform(0.05, 1.0) * t + rng.uniform(0, 2 * np.pi))
        c[k] *= amax / np.abs(c[k]).max()
    return c


def dim8_gue(seed=20260001, pts=4096):
    rng = np.random.default_rng(seed)
    return gue(rng, 8), [gue(rng, 8)], rng.uniform(-1, 1, (1, pts)), 1e-3, "none"


def heisenberg6(seed=20260002, pts=65536):
    rng = np.random.default_rng(seed)
    n = 6
    j = rng.uniform(0.5, 1.5, n - 1)
    h0 = sum(j[i] * sum(site_op(s, i, n) @ site_op(s, i + 1, n) for s in (SX, SY, SZ)) for i in range(n - 1))
    hk = [sum(site_op(s, i, n) for i in range(n)) for s in (SX, SY, SZ)]
    dt = 5e-3
    return h0, hk, sum_of_sines(rng, 3, pts, dt), dt, "simpson"


def ising8(seed=20260003, pts=16384):
    rng = np.random.default_rng(seed)
    n = 8
    h = rng.uniform(0, 2, n)
    h0 = sum(site_op(SZ, i, n) @ site_op(SZ, i + 1, n) for i in range(n - 1)) + sum(h[i] * site_op(SX, i, n) for i in range(n))
    hk = [site_op(SX, i, n) for i in range(n)]
    raw = rng.normal(size=(n, pts + 64))
    w = np.exp(-0.5 * (np.arange(-32, 33) / 8.0) ** 2)
    w /= w.sum()
    c = np.stack([np.convolve(r, w, mode="valid")[:pts] for r in raw]) * 4.0
    return h0, hk, c, 2e-3, "none"


def gue32_grape(seed=20260004, pts=4096):
    rng = np.random.default_rng(seed)
    h = [gue(rng, 32) for _ in range(5)]
    c = rng.uniform(-1, 1, (4, pts))
    return h[0], h[1:], c, 2e-3, "none", haar_unitary(rng, 32)


def ensemble512(seed=20260005, count=512, dt=1e-2):
    """512 independent systems: dim = round(exp(U[ln 4, ln 128])), 1-4 controls, 256-8192 steps,
    GUE H0 and controls (spectral norm 1), controls uniform in [-1, 1], no quadrature."""
    rng = np.random.default_rng(seed)
    dims = np.rint(np.exp(rng.uniform(np.log(4), np.log(128), count))).astype(int)
    amps = rng.integers(1, 5, count)
    pts = rng.integers(256, 8193, count)
    hs = [[gue(rng, int(d)) for _ in range(int(a) + 1)] for d, a in zip(dims, amps)]
    cs = [rng.uniform(-1, 1, (int(a), int(p))) for a, p in zip(amps, pts)]
    return dims, amps, pts, dt, hs, cs
