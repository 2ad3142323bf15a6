"""CPU oracle for Parament's equiprop path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker or the timed CPU
comparator.  The shipped library never calls it.

It restates, in numpy FP64, the algorithm of the reference implementation
(parament-integrator/parament, ``src/cuda``):

* ``one_norm``                   -- mathhelper.cpp:82-114 (max row sum of |M|,
                                    row-major input)
* ``select_iteration_cycles_*``  -- parament.cpp:723-766 (threshold tables)
* ``bessel_coefficients``        -- parament.cpp:403-407, mathhelper.cpp:35-80
                                    (a_k = (-i)^k J_k(tau*Hnorm), glibc ``jn``)
* ``quadrature``                 -- control_expansion.cu:105-170 and
                                    parament.cpp:800-802, 820-831
* ``magnus_coefficients``        -- control_expansion.cu:27-60, commutator slots
                                    parament.cpp:289-359
* ``assemble``                   -- parament.cpp:486-563
* ``clenshaw``                   -- parament.cpp:569-652 (+ diagonal_add.cu:21-41)
* ``pairwise_tree``              -- parament.cpp:657-718
* ``equiprop``                   -- parament.cpp:791-851 (orchestration)

The reference works on column-major views of the row-major arrays the Python
wrapper passes (parament.py:197-198); every step commutes with transposition
except matrix products, whose order flips.  This restatement works directly in
the row-major view, so its tree multiplies later slices on the LEFT:
``R'_i = R_{2i+1} @ R_{2i}``.

How each function is pinned (tests/test_oracle.py): the M tables by the
reference's constants; the Bessel coefficients against scipy ``jv`` and the
identities J_0 + 2 sum J_2k = 1 and the generating function; Clenshaw against
the textbook Chebyshev series of exp(-i x t) on scalars and against
``scipy.linalg.expm`` (the reference's own test thresholds,
test_numerics.py:32-97); the tree against brute-force sequential products and
the reference's 10-step two-field test (test_numerics.py:100-117); the
quadrature rules by closed forms (Simpson of constant controls equals 'none' at
2 dt, midpoint of linear controls is the exact average); the docstring matrix
of parament.py:61-68.  Magnus (the reference ships no Magnus test): with
controls linear across every slice, H(t) = P + t Q, the step must equal
exp(Omega1 + Omega2) with the closed forms Omega1 = -i tau (H(0) + H(tau))/2 and
Omega2 = (tau^3/12)[P, Q], slice by slice to 1e-12; its global error must fall
16x per halving of dt against a fine Gauss-Legendre Magnus product (Simpson
means alone: 4x); with all [H_a, H_b] = 0 it reduces to Simpson.  A flipped
sign or a wrong factor on either commutator row fails the first two.
"""
from __future__ import annotations

import ctypes
import ctypes.util

import numpy as np

QUAD_NONE = 0x00000000
QUAD_MIDPOINT = 0x01000000
QUAD_SIMPSON = 0x02000000

_MODE_BY_NAME = {"none": QUAD_NONE, "midpoint": QUAD_MIDPOINT, "simpson": QUAD_SIMPSON}

# parament.cpp:725-738 (FP32) and :746-765 (FP64): (upper bound on Hnorm*tau, M)
FP32_TABLE = (
    (0.032516793, 3), (0.219062571, 5), (0.619625593, 7), (1.218059203, 9),
    (1.979888284, 11), (2.873301187, 13), (3.872963682, 15), (4.959398466, 17),
    (6.117657121, 19), (7.336154907, 21), (8.605792444, 23), (9.919320831, 25),
    (11.27088616, 27), (12.65570085, 29),
)
FP64_TABLE = (
    (0.000213616, 3), (0.00768149, 5), (0.0501474, 7), (0.162592, 9),
    (0.368382, 11), (0.676861, 13), (1.08784, 15), (1.59605, 17),
    (2.19402, 19), (2.87366, 21), (3.62716, 23), (4.44725, 25),
    (5.3274, 27), (6.26179, 29), (7.2453, 31), (8.27338, 33),
    (9.34206, 35), (10.4478, 37), (11.5875, 39), (12.7584, 41),
)

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.jn.restype = ctypes.c_double
_libm.jn.argtypes = (ctypes.c_int, ctypes.c_double)


def jn(k: int, x: float) -> float:
    """glibc ``jn`` -- the same routine the reference calls (mathhelper.cpp:61-67)."""
    return float(_libm.jn(int(k), float(x)))


def _select(table, hnorm: float, dt: float) -> int:
    x = hnorm * dt
    for bound, m in table:
        if x <= bound:
            return m
    return -1


def select_iteration_cycles_fp32(hnorm: float, dt: float) -> int:
    """parament.cpp:723-739."""
    return _select(FP32_TABLE, hnorm, dt)


def select_iteration_cycles_fp64(hnorm: float, dt: float) -> int:
    """parament.cpp:744-766."""
    return _select(FP64_TABLE, hnorm, dt)


def chebyshev_tail(m: int, x: float) -> float:
    """2 sum_{k>m} |J_k(x)|: spectral-norm bound of the truncation error of one slice."""
    return 2.0 * sum(abs(jn(k, x)) for k in range(m + 1, m + 41))


def select_iteration_cycles_mixed(hnorm: float, tau: float, n_slices: int) -> int:
    """Order of this build's mixed-precision path (no reference counterpart): the FP32 table's
    M, raised by 2 until n_slices * chebyshev_tail(M) <= 2e-6 (systematic truncation error
    summed over the product)."""
    m = select_iteration_cycles_fp32(hnorm, tau)
    if m < 3:
        return m
    x = hnorm * tau
    while m + 2 < 48 and n_slices * chebyshev_tail(m, x) > 2e-6:
        m += 2
    return m


def one_norm(m: np.ndarray) -> float:
    """Max over rows of the summed moduli of a row-major matrix (mathhelper.cpp:82-114).

    For complex64 input the reference takes the modulus in single precision and
    sums into a double; ``np.abs`` of a complex64 array does the same.
    """
    return float(np.max(np.sum(np.abs(m).astype(np.float64), axis=1)))


def hamiltonian_norm(h0: np.ndarray, hks) -> float:
    """Hnorm = ||H0|| + sum_k ||H_k|| over the control Hamiltonians set (parament.cpp:280-284)."""
    return one_norm(h0) + sum(one_norm(h) for h in hks)


def bessel_coefficients(m: int, x: float) -> np.ndarray:
    """a_k = (-i)^k J_k(x) for k = 0..m (parament.cpp:403-407, mathhelper.cpp:35-80)."""
    powers = (1.0, -1j, -1.0, 1j)
    return np.array([powers[k % 4] * jn(k, x) for k in range(m + 1)], dtype=np.complex128)


def slice_count(pts: int, mode: int, use_magnus: bool = False) -> int:
    """Number of propagator slices per mode (parament.cpp:820-831)."""
    if use_magnus or mode == QUAD_SIMPSON:
        return max(pts - 1, 0) // 2
    if mode == QUAD_MIDPOINT:
        return max(pts - 1, 0)
    return pts


def quadrature(c: np.ndarray, mode: int) -> np.ndarray:
    """Per-slice control coefficients c'[k, j] (control_expansion.cu:105-170).

    c has shape (amps, pts).  Simpson uses ((c[2i] + 4 c[2i+1]) + c[2i+2]) / 6
    and silently drops the last sample for even pts (control_expansion.cu:144-146).
    """
    c = np.asarray(c, dtype=np.complex128)
    if mode == QUAD_NONE:
        return c.copy()
    if mode == QUAD_MIDPOINT:
        return 0.5 * (c[:, :-1] + c[:, 1:])
    if mode == QUAD_SIMPSON:
        n = slice_count(c.shape[1], mode)
        return ((c[:, 0:2 * n:2] + 4.0 * c[:, 1:2 * n:2]) + c[:, 2:2 * n + 1:2]) / 6.0
    raise ValueError("unknown quadrature mode")


def pair_index(j: int, k: int, amps: int) -> int:
    """Slot of the pairwise commutator [H_j, H_k], j < k, among the amps*(amps-1)/2 pairs.

    The reference uses slot j + k - 1 (parament.cpp:330, control_expansion.cu:53),
    which is unique only for amps <= 3; this lexicographic index agrees with it
    there and stays unique for amps >= 4.
    """
    return j * amps - j * (j + 1) // 2 + (k - j - 1)


def magnus_operators(h0: np.ndarray, hks) -> list:
    """[H0, H_k] for each k, then [H_j, H_k] for j < k (row-major view of parament.cpp:295-358)."""
    amps = len(hks)
    ops = [h0 @ hk - hk @ h0 for hk in hks]
    pairs = [None] * (amps * (amps - 1) // 2)
    for j in range(amps):
        for k in range(j + 1, amps):
            pairs[pair_index(j, k, amps)] = hks[j] @ hks[k] - hks[k] @ hks[j]
    return ops + pairs


def magnus_coefficients(c: np.ndarray, tau: float) -> np.ndarray:
    """Rows: Simpson means, then i*tau/12*(c_k[2i+2]-c_k[2i]), then pair terms.

    control_expansion.cu:27-60 with tau = 2 dt (parament.cpp:800-802).
    """
    c = np.asarray(c, dtype=np.complex128)
    amps, pts = c.shape
    n = slice_count(pts, QUAD_SIMPSON, True)
    fac = 1j * tau / 12.0
    mean = ((c[:, 0:2 * n:2] + 4.0 * c[:, 1:2 * n:2]) + c[:, 2:2 * n + 1:2]) / 6.0
    grad = (c[:, 2:2 * n + 1:2] - c[:, 0:2 * n:2]) * fac
    pairs = np.zeros((amps * (amps - 1) // 2, n), dtype=np.complex128)
    for j in range(amps):
        for k in range(j + 1, amps):
            pairs[pair_index(j, k, amps)] = (c[j, 0:2 * n:2] * c[k, 2:2 * n + 1:2]
                                             - c[j, 2:2 * n + 1:2] * c[k, 0:2 * n:2]) * fac
    return np.concatenate([mean, grad, pairs], axis=0)


def assemble(h0: np.ndarray, ops, coeffs: np.ndarray, j0: int = 0, j1: int | None = None) -> np.ndarray:
    """X_j = H0 + sum_k coeffs[k, j] ops[k] for slices j0..j1 (parament.cpp:486-563)."""
    j1 = coeffs.shape[1] if j1 is None else j1
    x = np.broadcast_to(np.asarray(h0, np.complex128), (j1 - j0,) + h0.shape).copy()
    for k, op in enumerate(ops):
        x += coeffs[k, j0:j1, None, None] * np.asarray(op, np.complex128)[None]
    return x


def clenshaw(x: np.ndarray, a: np.ndarray, hnorm: float) -> np.ndarray:
    """Batched Clenshaw sum of the Chebyshev series (parament.cpp:569-652).

    With s = 2/Hnorm: b_M = a_M I, b_{M-1} = s X b_M + a_{M-1} I,
    b_k = s X b_{k+1} - b_{k+2} + a_k I, and U = s X b_1 - 2 b_2 + a_0 I,
    i.e. U = sum_k (2 - delta_k0) a_k T_k(X / Hnorm).  M = len(a) - 1 must be
    odd and >= 3, as the reference's loop requires.
    """
    m = len(a) - 1
    if m < 3 or m % 2 == 0:
        raise ValueError("Chebyshev order must be odd and >= 3")
    d = x.shape[-1]
    eye = np.eye(d, dtype=np.complex128)
    s = 2.0 / hnorm
    b2 = np.broadcast_to(a[m] * eye, x.shape)
    b1 = s * (x @ b2) + a[m - 1] * eye
    for k in range(m - 2, 0, -1):
        b1, b2 = s * (x @ b1) - b2 + a[k] * eye, b1
    return s * (x @ b1) - 2.0 * b2 + a[0] * eye


def pairwise_tree(r: np.ndarray) -> np.ndarray:
    """Log-depth product with odd carry (parament.cpp:657-718), later slices on the left."""
    r = np.asarray(r)
    while r.shape[0] > 1:
        n = r.shape[0]
        h = n // 2
        w = r[1:2 * h:2] @ r[0:2 * h:2]
        if n % 2:
            w = np.concatenate([w, r[n - 1:n]], axis=0)
        r = w
    return r[0]


def equiprop(h0, hks, carr, dt: float, mode=QUAD_NONE, use_magnus: bool = False,
             cycles: int | None = None, precision: str = "fp64", amps: int | None = None,
             chunk: int = 4096) -> np.ndarray:
    """Propagator U = U_{N-1} ... U_0 exactly as Parament_equiprop defines it (parament.cpp:791-851).

    ``precision`` only selects the iteration table ('fp32' or 'fp64'); all
    arithmetic is FP64.  ``cycles`` forces a manual M.  ``carr`` has shape
    (amps, pts); fewer rows than control Hamiltonians means the rest have zero
    amplitude (parament.h:228-230).
    """
    if isinstance(mode, str):
        mode = _MODE_BY_NAME[mode]
    h0 = np.asarray(h0, np.complex128)
    hks = [np.asarray(h, np.complex128) for h in hks]
    carr = np.atleast_2d(np.asarray(carr, np.complex128))
    if amps is None:
        amps = carr.shape[0]
    hnorm = hamiltonian_norm(h0, hks)
    tau = 2.0 * dt if (use_magnus or mode == QUAD_SIMPSON) else dt
    if cycles is None:
        table = FP32_TABLE if precision == "fp32" else FP64_TABLE
        m = _select(table, hnorm, tau)
        if m < 3:
            raise RuntimeError("Timestep too large")
    else:
        m = int(cycles)
    a = bessel_coefficients(m, tau * hnorm)
    c = np.zeros((len(hks), carr.shape[1]), np.complex128)
    c[:amps] = carr[:amps]
    if use_magnus:
        coeffs = magnus_coefficients(c, tau)
        ops = list(hks) + magnus_operators(h0, hks)
    else:
        coeffs = quadrature(c, mode)
        ops = hks
    n = coeffs.shape[1]
    if n == 0:
        raise ValueError("no slices")
    parts = []
    for j0 in range(0, n, chunk):
        j1 = min(n, j0 + chunk)
        parts.append(clenshaw(assemble(h0, ops, coeffs, j0, j1), a, hnorm))
    return pairwise_tree(np.concatenate(parts, axis=0))


def slice_hamiltonians(h0, hks, carr, dt: float, mode=QUAD_NONE, use_magnus: bool = False):
    """Per-slice effective Hamiltonians X_j and the slice length tau (for the expm oracle)."""
    if isinstance(mode, str):
        mode = _MODE_BY_NAME[mode]
    h0 = np.asarray(h0, np.complex128)
    hks = [np.asarray(h, np.complex128) for h in hks]
    carr = np.atleast_2d(np.asarray(carr, np.complex128))
    c = np.zeros((len(hks), carr.shape[1]), np.complex128)
    c[:carr.shape[0]] = carr
    tau = 2.0 * dt if (use_magnus or mode == QUAD_SIMPSON) else dt
    if use_magnus:
        coeffs = magnus_coefficients(c, tau)
        ops = list(hks) + magnus_operators(h0, hks)
    else:
        coeffs = quadrature(c, mode)
        ops = hks
    return assemble(h0, ops, coeffs), tau
