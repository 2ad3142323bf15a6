"""Stepwise scipy-expm propagator -- TEST INFRASTRUCTURE ONLY (the "truth" oracle).

U = expm(-i tau X_{N-1}) ... expm(-i tau X_0) in complex128, with the same
quadrature-averaged per-slice Hamiltonians X_j that the reference builds
(parament.cpp:486-563).  This is the comparison the reference's own tests make
(test_numerics.py:96, 166-171).  Only tests/, smoke() and bench.py's
cpu_baseline leg may use it.

``expm_stepwise_parallel`` computes the same product on several host processes: each takes a
contiguous block of slices, forms their Hamiltonians from its share of the samples, exponentiates
them with one BLAS thread and returns the block's ordered product; the blocks are multiplied in
time order (later blocks on the left).  Only the association of the matrix products differs from
the sequential form (rounding-level differences, ~1e-15 relative).
"""
from __future__ import annotations

import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np
import scipy.linalg

from .equiprop_oracle import QUAD_MIDPOINT, QUAD_NONE, _MODE_BY_NAME, slice_count, slice_hamiltonians


def expm_stepwise(h0, hks, carr, dt: float, mode="none", use_magnus: bool = False,
                  xs=None, tau=None) -> np.ndarray:
    from threadpoolctl import threadpool_limits
    if xs is None:
        xs, tau = slice_hamiltonians(h0, hks, carr, dt, mode, use_magnus)
    d = xs.shape[-1]
    u = np.eye(d, dtype=np.complex128)
    # one BLAS thread: the products are small, and threaded BLAS spends far longer waking its
    # threads than computing them
    with threadpool_limits(1):
        for x in xs:
            u = scipy.linalg.expm(-1j * tau * x) @ u
    return u


def _sample_window(j0: int, j1: int, mode: int, use_magnus: bool):
    """Samples [s0, s1) that slices j0 .. j1 - 1 read (parament.cpp:820-831 slice layout)."""
    if use_magnus or mode not in (QUAD_NONE, QUAD_MIDPOINT):
        return 2 * j0, 2 * j1 + 1
    if mode == QUAD_MIDPOINT:
        return j0, j1 + 1
    return j0, j1


def _block_product(args):
    return expm_stepwise(*args)


def expm_stepwise_parallel(h0, hks, carr, dt: float, mode="none", use_magnus: bool = False,
                           workers: int | None = None) -> np.ndarray:
    if isinstance(mode, str):
        mode = _MODE_BY_NAME[mode]
    carr = np.atleast_2d(np.asarray(carr))
    n = slice_count(carr.shape[1], mode, use_magnus)
    workers = max(1, min(workers or os.cpu_count() or 1, n))
    bounds = np.linspace(0, n, workers + 1).astype(int)
    jobs = []
    for j0, j1 in zip(bounds[:-1], bounds[1:]):
        if j1 > j0:
            s0, s1 = _sample_window(int(j0), int(j1), mode, use_magnus)
            jobs.append((h0, hks, carr[:, s0:s1], dt, mode, use_magnus))
    with ProcessPoolExecutor(max_workers=len(jobs)) as ex:
        parts = list(ex.map(_block_product, jobs))
    u = parts[0]
    for p in parts[1:]:
        u = p @ u
    return u
