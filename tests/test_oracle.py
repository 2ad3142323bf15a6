"""Pins the CPU oracle (oracle/equiprop_oracle.py) against the reference's known answers
and against what the mathematics fixes.  CPU only."""
import numpy as np
import pytest
import scipy.linalg
from scipy.special import jv

from oracle import equiprop_oracle as eo
from oracle.expm_stepwise import expm_stepwise
from tests.helpers import gue, random_hamiltonians, rel_err


# ---------------- OneNorm and Hnorm (mathhelper.cpp:82-114, parament.cpp:280-284) ----------------

def test_one_norm_is_the_max_row_sum():
    """A non-Hermitian case whose max row sum (5), max column sum (7), Frobenius norm (5.48) and
    spectral norm (5.40) all differ: the reference sums the moduli along each row of the
    row-major input (mathhelper.cpp:88-96)."""
    m = np.array([[1, 2j], [0, 3 + 4j]])
    assert eo.one_norm(m) == 5.0
    assert eo.one_norm(m.T) == 7.0   # the transpose swaps rows and columns
    assert eo.one_norm(m.astype(np.complex64)) == 5.0


@pytest.mark.parametrize("seed", range(4))
def test_one_norm_equals_infinity_norm(seed):
    """Against numpy's induced infinity norm on random non-Hermitian complex matrices of odd and
    even dims; complex64 input takes each modulus in single precision (within 6e-8 of it)."""
    rng = np.random.default_rng(seed)
    d = 3 + 5 * seed
    m = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)) * (1 + seed)
    assert eo.one_norm(m) == pytest.approx(np.linalg.norm(m, np.inf), rel=1e-15)
    assert abs(eo.one_norm(m) - np.linalg.norm(m, 1)) > 1e-6   # not the column sums
    m32 = m.astype(np.complex64)
    assert eo.one_norm(m32) == pytest.approx(np.linalg.norm(m32.astype(np.complex128), np.inf), rel=1e-7)


def test_hamiltonian_norm_sums_the_operators():
    """Hnorm = ||H0|| + sum_k ||H_k||, each a max row sum: 5 + 7 + 2."""
    m = np.array([[1, 2j], [0, 3 + 4j]])
    assert eo.hamiltonian_norm(m, [m.T, 2 * np.eye(2)]) == 14.0
    # the reference docstring case: |diag(1, -1)| + |sigma_x| = 2
    assert eo.hamiltonian_norm(np.diag([1.0, -1.0]), [np.array([[0, 1], [1, 0]])]) == 2.0


# ---------------- iteration tables (parament.cpp:723-766) ----------------

@pytest.mark.parametrize("table,eps", [(eo.FP32_TABLE, np.finfo(np.float32).eps),
                                       (eo.FP64_TABLE, np.finfo(np.float64).eps)])
def test_tables_follow_truncation_rule(table, eps):
    """Each threshold x_M is where the Chebyshev tail 2 sum_{k>M} |J_k(x)| reaches ~0.02-0.05 eps,
    M runs over consecutive odd numbers from 3 and the thresholds increase: a wrong digit in
    a threshold moves its tail out of the band."""
    ms = [m for _, m in table]
    assert ms == list(range(3, 3 + 2 * len(ms), 2))
    xs = [x for x, _ in table]
    assert all(a < b for a, b in zip(xs, xs[1:]))
    for x, m in table:
        tail = 2 * sum(abs(jv(k, x)) for k in range(m + 1, m + 60))
        assert 0.015 * eps < tail < 0.06 * eps, (m, x, tail / eps)
        prev = 2 * sum(abs(jv(k, x)) for k in range(m - 1, m + 60))
        assert prev > 0.4 * eps  # M - 2 would not suffice


def test_table_boundaries():
    assert eo.select_iteration_cycles_fp32(1.0, 0.032516793) == 3
    assert eo.select_iteration_cycles_fp32(1.0, 0.0325168) == 5
    assert eo.select_iteration_cycles_fp32(2.0, 12.65570085 / 2) == 29
    assert eo.select_iteration_cycles_fp32(1.0, 12.66) == -1
    assert eo.select_iteration_cycles_fp64(1.0, 0.000213616) == 3
    assert eo.select_iteration_cycles_fp64(1.0, 12.7584) == 41
    assert eo.select_iteration_cycles_fp64(1.0, 12.76) == -1


# ---------------- Bessel coefficients (mathhelper.cpp:35-80) ----------------

@pytest.mark.parametrize("x", [1e-4, 0.0325, 0.3376, 1.0, 5.0, 12.7])
def test_bessel_against_scipy_and_identities(x):
    a = eo.bessel_coefficients(41, x)
    k = np.arange(42)
    assert np.allclose(np.abs(a), np.abs(jv(k, x)), rtol=1e-13, atol=1e-300)
    # (-i)^k phase pattern
    assert np.allclose(a[0::4].imag, 0) and np.allclose(a[1::4].real, 0)
    # J_0 + 2 sum_{k>=1} J_{2k} = 1
    j = jv(np.arange(80), x)
    assert abs(j[0] + 2 * j[2::2].sum() - 1) < 1e-14
    # generating function: exp(-i x cos t) = J_0 + 2 sum (-i)^k J_k cos(k t)
    t = 0.7
    series = eo.bessel_coefficients(60, x) * np.r_[1.0, 2 * np.cos(np.arange(1, 61) * t)]
    assert abs(series.sum() - np.exp(-1j * x * np.cos(t))) < 1e-13


# ---------------- Clenshaw (parament.cpp:569-652) ----------------

def test_clenshaw_scalar_equals_chebyshev_series():
    """On 1x1 'matrices' the recurrence must equal sum (2 - delta_k0) a_k T_k(y) exactly."""
    hnorm, x = 2.5, 0.9
    a = eo.bessel_coefficients(9, x)
    ys = np.linspace(-1, 1, 11)
    xs = (ys * hnorm).reshape(-1, 1, 1).astype(complex)
    u = eo.clenshaw(xs, a, hnorm)[:, 0, 0]
    t = np.cos(np.outer(np.arccos(ys), np.arange(10)))
    direct = t @ (a * np.r_[1.0, 2 * np.ones(9)])
    assert np.allclose(u, direct, atol=1e-14)
    # and at this order the series is exp(-i x y) to truncation accuracy
    assert np.allclose(u, np.exp(-1j * x * ys), atol=1e-8)


@pytest.mark.parametrize("dim", [2, 4, 16])
def test_clenshaw_matches_expm_reference_thresholds(dim):
    """The reference's single-step test (test_numerics.py:32-97): sum |U - expm| < eps * d^2."""
    h0, h1 = random_hamiltonians(dim)
    for prec, eps in (("fp64", np.finfo(np.float64).eps), ("fp32", np.finfo(np.float32).eps)):
        u = eo.equiprop(h0, [h1], np.zeros((1, 1)), 0.01, precision=prec)
        assert np.abs(u - scipy.linalg.expm(-1j * 0.01 * h0)).sum() < eps * dim * dim


def test_clenshaw_rejects_even_or_small_orders():
    x = np.eye(2, dtype=complex)[None]
    for m in (1, 2, 4):
        with pytest.raises(ValueError):
            eo.clenshaw(x, eo.bessel_coefficients(m, 0.1), 1.0)


def test_debug_expm_trick(golden):
    """parament.expm(m): H0 = 1j m, H1 = m, c = 0, dt = 1 gives exp(m) (debug_functions.py:22-31)."""
    for dim in (2, 4):
        h0 = golden[f"seed27_d{dim}_h0"]
        u = eo.equiprop(1j * h0, [h0], np.zeros((1, 1)), 1.0, precision="fp32")
        assert np.abs(u - golden[f"seed27_d{dim}_expm_debug"]).sum() < np.finfo(np.float32).eps * dim * dim


def test_docstring_matrix(golden):
    u = eo.equiprop(golden["doc_h0"], [golden["doc_h1"]], np.zeros((1, 1)), 1.0, precision="fp32")
    assert np.allclose(u, golden["doc_expected"], atol=1e-8)


# ---------------- product tree (parament.cpp:657-718) ----------------

@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 13, 17, 33])
def test_tree_equals_time_ordered_product(n):
    rng = np.random.default_rng(n)
    r = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    seq = np.eye(3, dtype=complex)
    for j in range(n):
        seq = r[j] @ seq
    assert rel_err(eo.pairwise_tree(r), seq) < 1e-12


def test_tree_association_matches_aligned_blocks():
    """Node (l, i) of the pairing covers slices [i 2^l, min((i+1) 2^l, N)): the tree of a
    length-N list equals the tree of its aligned 8-blocks' trees, bit for bit."""
    rng = np.random.default_rng(3)
    n = 45
    r = rng.normal(size=(n, 4, 4)) + 1j * rng.normal(size=(n, 4, 4))
    blocks = np.stack([eo.pairwise_tree(r[i:i + 8]) for i in range(0, n, 8)])
    assert np.array_equal(eo.pairwise_tree(blocks), eo.pairwise_tree(r))


@pytest.mark.parametrize("dim", [2, 16])
def test_time_ordering_reference_case(golden, dim):
    """The reference's only ordering test (test_numerics.py:100-117), threshold 1e-6."""
    h0, h1 = random_hamiltonians(dim)
    c = golden[f"multi_d{dim}_c"]
    u = eo.equiprop(h0, [h1, h0], c, 0.01, precision="fp32")
    assert np.linalg.norm(u - golden[f"multi_d{dim}_expected"]) < 1e-6


# ---------------- quadrature (control_expansion.cu:105-170) ----------------

def test_slice_counts():
    assert eo.slice_count(10, eo.QUAD_NONE) == 10
    assert eo.slice_count(10, eo.QUAD_MIDPOINT) == 9
    assert eo.slice_count(10, eo.QUAD_SIMPSON) == 4   # even pts drops the last sample
    assert eo.slice_count(11, eo.QUAD_SIMPSON) == 5
    assert eo.slice_count(2, eo.QUAD_SIMPSON) == 0
    assert eo.slice_count(1, eo.QUAD_MIDPOINT) == 0


def test_quadrature_closed_forms():
    t = np.arange(9, dtype=float)
    lin = np.stack([2 * t + 1, -t])
    mid = eo.quadrature(lin, eo.QUAD_MIDPOINT)
    assert np.allclose(mid, np.stack([2 * (t[:-1] + 0.5) + 1, -(t[:-1] + 0.5)]))
    quad = np.stack([t ** 2])
    simp = eo.quadrature(quad, eo.QUAD_SIMPSON)
    # Simpson is exact for quadratics: mean over [2j, 2j+2] of t^2
    exact = np.array([((2 * j + 2) ** 3 - (2 * j) ** 3) / 6 for j in range(4)])
    assert np.allclose(simp[0], exact)
    # even pts drops the last sample
    assert eo.quadrature(np.ones((1, 10)), eo.QUAD_SIMPSON).shape == (1, 4)


def test_simpson_constant_equals_none_at_twice_dt():
    rng = np.random.default_rng(11)
    h = [gue(rng, 4) for _ in range(3)]
    c = np.array([[0.3], [-0.7]])
    u_s = eo.equiprop(h[0], h[1:], np.repeat(c, 9, axis=1), 0.05, mode="simpson")
    u_n = eo.equiprop(h[0], h[1:], np.repeat(c, 4, axis=1), 0.1, mode="none")
    assert rel_err(u_s, u_n) < 1e-13


# ---------------- Magnus (control_expansion.cu:27-60, parament.cpp:289-359) ----------------

def test_pair_index_matches_reference_slots_for_three_controls():
    # reference slot j + k - 1 for j < k (unique for amps <= 3)
    for (j, k) in ((0, 1), (0, 2), (1, 2)):
        assert eo.pair_index(j, k, 3) == j + k - 1
    idx = sorted(eo.pair_index(j, k, 5) for j in range(5) for k in range(j + 1, 5))
    assert idx == list(range(10))


def test_magnus_commuting_case_reduces_to_simpson():
    rng = np.random.default_rng(5)
    q = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))[0]
    diag = lambda: q @ np.diag(rng.normal(size=4)) @ q.conj().T
    h = [diag() for _ in range(3)]
    c = rng.uniform(-1, 1, (2, 17))
    u_m = eo.equiprop(h[0], h[1:], c, 0.05, mode="simpson", use_magnus=True)
    u_s = eo.equiprop(h[0], h[1:], c, 0.05, mode="simpson")
    assert rel_err(u_m, u_s) < 1e-12


def test_magnus_is_more_accurate_than_simpson():
    rng = np.random.default_rng(1)
    h0, h1, h2 = gue(rng, 4), gue(rng, 4), gue(rng, 4)
    f = lambda t: np.stack([np.sin(3 * t + 0.3), np.cos(2 * t)])
    dt, pts = 0.05, 41
    c = f(np.arange(pts) * dt)
    fine = 64
    u = np.eye(4, dtype=complex)
    for s in range((pts - 1) * fine):
        cc = f((s + 0.5) * dt / fine)
        u = scipy.linalg.expm(-1j * dt / fine * (h0 + cc[0] * h1 + cc[1] * h2)) @ u
    e_s = np.linalg.norm(eo.equiprop(h0, [h1, h2], c, dt, mode="simpson") - u)
    e_m = np.linalg.norm(eo.equiprop(h0, [h1, h2], c, dt, mode="simpson", use_magnus=True) - u)
    assert e_m < 0.5 * e_s


def test_magnus_equals_closed_form_for_piecewise_linear_controls():
    # With H(t) = P + t Q on a slice of length tau, the Magnus series stops after two terms up to
    # O(tau^5), and those two are closed forms: Omega1 = -i tau (H(0) + H(tau)) / 2 and
    # Omega2 = -1/2 int_0^tau int_0^t1 [H(t1), H(t2)] = (tau^3 / 12) [P, Q].  Controls that are
    # linear across every slice (mid sample = mean of the ends) make the fourth-order step exact
    # in those terms, so the oracle must equal the product of exp(Omega1 + Omega2) slice by slice.
    rng = np.random.default_rng(12)
    d, amps, n, dt = 4, 3, 5, 0.07
    h0 = gue(rng, d)
    hks = [gue(rng, d) for _ in range(amps)]
    c = np.empty((amps, 2 * n + 1))
    c[:, 0::2] = rng.uniform(-1.5, 1.5, (amps, n + 1))
    c[:, 1::2] = 0.5 * (c[:, 0:-1:2] + c[:, 2::2])
    tau = 2 * dt
    ham = lambda col: h0 + sum(col[k] * hks[k] for k in range(amps))
    u = np.eye(d, dtype=complex)
    for j in range(n):
        ha, hb = ham(c[:, 2 * j]), ham(c[:, 2 * j + 2])
        p, q = ha, (hb - ha) / tau
        omega = -1j * tau * 0.5 * (ha + hb) + tau ** 3 / 12 * (p @ q - q @ p)
        u = scipy.linalg.expm(omega) @ u
    got = eo.equiprop(h0, hks, c, dt, mode="simpson", use_magnus=True)
    assert rel_err(got, u) < 1e-12
    # the commutator terms are not negligible here: Simpson alone misses by far more
    assert rel_err(eo.equiprop(h0, hks, c, dt, mode="simpson"), u) > 1e-4


def test_magnus_converges_at_fourth_order():
    # Global error of the fourth-order Magnus step falls by 2^4 per halving of dt, Simpson
    # means alone (exp(-i tau Hbar)) by 2^2.  A wrong factor or sign on either commutator row
    # leaves an O(tau^3) local error and shows as second order.
    rng = np.random.default_rng(3)
    h0, h1, h2 = gue(rng, 3), gue(rng, 3), gue(rng, 3)
    f = lambda t: np.stack([np.sin(2 * t + 0.4), np.cos(3 * t) ** 2])
    t_end = 0.8

    def truth():
        fine = 4096
        h = t_end / fine
        u = np.eye(3, dtype=complex)
        for s in range(fine):   # Gauss-Legendre 2-point Magnus step, order 4, on a fine grid
            ta, tb = (s + 0.5 - 0.5 / np.sqrt(3)) * h, (s + 0.5 + 0.5 / np.sqrt(3)) * h
            ca, cb = f(ta), f(tb)
            ha = h0 + ca[0] * h1 + ca[1] * h2
            hb = h0 + cb[0] * h1 + cb[1] * h2
            om = -1j * h * 0.5 * (ha + hb) + np.sqrt(3) / 12 * h * h * (ha @ hb - hb @ ha)
            u = scipy.linalg.expm(om) @ u
        return u

    u = truth()
    errs_m, errs_s = [], []
    for pts in (9, 17, 33):
        dt = t_end / (pts - 1)
        c = f(np.arange(pts) * dt)
        errs_m.append(np.linalg.norm(eo.equiprop(h0, [h1, h2], c, dt, mode="simpson", use_magnus=True) - u))
        errs_s.append(np.linalg.norm(eo.equiprop(h0, [h1, h2], c, dt, mode="simpson") - u))
    for a, b in zip(errs_m, errs_m[1:]):
        assert 12 < a / b < 20
    for a, b in zip(errs_s, errs_s[1:]):
        assert 3 < a / b < 5


# ---------------- whole pipeline vs the stepwise expm oracle ----------------

@pytest.mark.parametrize("mode,mag", [("none", False), ("midpoint", False), ("simpson", False), ("simpson", True)])
def test_equiprop_matches_stepwise_expm(mode, mag):
    rng = np.random.default_rng(2)
    h = [gue(rng, 6) for _ in range(3)]
    c = rng.uniform(-1, 1, (2, 33))
    u = eo.equiprop(h[0], h[1:], c, 0.02, mode=mode, use_magnus=mag)
    ref = expm_stepwise(h[0], h[1:], c, 0.02, mode=mode, use_magnus=mag)
    assert rel_err(u, ref) < 1e-13


def test_fewer_controls_than_hamiltonians_means_zero_amplitude():
    rng = np.random.default_rng(4)
    h = [gue(rng, 4) for _ in range(3)]
    c = rng.uniform(-1, 1, (1, 9))
    u1 = eo.equiprop(h[0], h[1:], c, 0.05, mode="simpson")
    u2 = eo.equiprop(h[0], h[1:], np.vstack([c, np.zeros((1, 9))]), 0.05, mode="simpson")
    assert np.array_equal(u1, u2)


def test_golden_regression_vectors(golden):
    hs, c = golden["mode_h"], golden["mode_c"]
    for name, mode, mag in (("none", "none", False), ("midpoint", "midpoint", False),
                            ("simpson", "simpson", False), ("magnus", "simpson", True)):
        for prec in ("fp32", "fp64"):
            u = eo.equiprop(hs[0], list(hs[1:]), c, 0.05, mode=mode, use_magnus=mag, precision=prec)
            assert rel_err(u, golden[f"mode_{name}_{prec}"]) < 1e-14


def test_mixed_order_bounds_accumulated_truncation():
    """The mixed path's order keeps N * tail(M) <= 2e-6 and never goes below the FP32 table."""
    for x, n in ((0.0308, 4096), (0.3376, 32767), (0.3376, 10), (1.5, 100000), (0.0005, 1)):
        m32 = eo.select_iteration_cycles_fp32(x, 1.0)
        m = eo.select_iteration_cycles_mixed(x, 1.0, n)
        assert m >= m32 and m % 2 == 1
        assert n * eo.chebyshev_tail(m, x) <= 2e-6
        if m > m32:
            assert n * eo.chebyshev_tail(m - 2, x) > 2e-6
    # the case that motivated it: dim-32 GUE sum, 4096 slices at x = 0.0308 -> M = 5 (FP32 table: 3)
    assert eo.select_iteration_cycles_mixed(0.0308, 1.0, 4096) == 5
    assert eo.select_iteration_cycles_mixed(0.3376, 1.0, 32767) == 7


# ---------------- gradient oracle (oracle/grape_oracle.py) ----------------

def test_frechet_block_matches_scipy_expm_frechet():
    from oracle.grape_oracle import frechet_block
    rng = np.random.default_rng(9)
    a = -1j * 0.3 * gue(rng, 5)
    e = -1j * 0.3 * gue(rng, 5)
    ref = scipy.linalg.expm_frechet(a, e, compute_expm=False)
    assert np.allclose(frechet_block(a, e), ref, atol=1e-13)


@pytest.mark.parametrize("mode", ["none", "midpoint", "simpson"])
def test_grape_oracle_matches_central_differences(mode):
    from oracle.grape_oracle import fidelity, fidelity_gradient
    from parament_amd.workloads import haar_unitary
    rng = np.random.default_rng(21)
    h = [gue(rng, 3) for _ in range(3)]
    c = rng.uniform(-1, 1, (2, 9))
    t = haar_unitary(rng, 3)
    f, grad = fidelity_gradient(h[0], h[1:], c, 0.2, mode, t)
    assert abs(f - fidelity(h[0], h[1:], c, 0.2, mode, t)) < 1e-13
    hstep = 1e-6
    for k in range(2):
        for p in range(9):
            cp, cm = c.copy(), c.copy()
            cp[k, p] += hstep
            cm[k, p] -= hstep
            fd = (fidelity(h[0], h[1:], cp, 0.2, mode, t) - fidelity(h[0], h[1:], cm, 0.2, mode, t)) / (2 * hstep)
            assert abs(fd - grad[k, p]) < 1e-8, (k, p, fd, grad[k, p])


def test_grape_oracle_magnus_matches_central_differences():
    """Magnus slices: the exact-derivative oracle against central differences of the stepwise
    exponential of the Magnus exponents, for every sample of three controls (pair terms
    included) and an odd and an even sample count."""
    from oracle.grape_oracle import fidelity, fidelity_gradient
    from parament_amd.workloads import haar_unitary
    rng = np.random.default_rng(23)
    h = [gue(rng, 3) for _ in range(4)]
    t = haar_unitary(rng, 3)
    for pts in (9, 8):
        c = rng.uniform(-1, 1, (3, pts))
        f, grad = fidelity_gradient(h[0], h[1:], c, 0.2, "simpson", t, use_magnus=True)
        assert abs(f - fidelity(h[0], h[1:], c, 0.2, "simpson", t, use_magnus=True)) < 1e-13
        hstep = 1e-6
        for k in range(3):
            for p in range(pts):
                cp, cm = c.copy(), c.copy()
                cp[k, p] += hstep
                cm[k, p] -= hstep
                fd = (fidelity(h[0], h[1:], cp, 0.2, "simpson", t, True)
                      - fidelity(h[0], h[1:], cm, 0.2, "simpson", t, True)) / (2 * hstep)
                assert abs(fd - grad[k, p]) < 1e-8, (pts, k, p, fd, grad[k, p])
        if pts % 2 == 0:
            assert np.all(grad[:, -1] == 0)   # the sample Simpson drops


def test_grape_oracle_window_is_the_restricted_sum():
    """The slice window sums exactly the chosen slices' contributions: two windows that split
    the slices add up to the full gradient."""
    from oracle.grape_oracle import fidelity_gradient
    from parament_amd.workloads import haar_unitary
    rng = np.random.default_rng(29)
    h = [gue(rng, 4) for _ in range(3)]
    c = rng.uniform(-1, 1, (2, 12))
    t = haar_unitary(rng, 4)
    for mode, magnus, n in (("none", False, 12), ("simpson", False, 5), ("simpson", True, 5)):
        _, full = fidelity_gradient(h[0], h[1:], c, 0.1, mode, t, use_magnus=magnus)
        _, a = fidelity_gradient(h[0], h[1:], c, 0.1, mode, t, use_magnus=magnus, slices=range(0, 3))
        _, b = fidelity_gradient(h[0], h[1:], c, 0.1, mode, t, use_magnus=magnus, slices=range(3, n))
        assert np.allclose(a + b, full, atol=1e-15)


@pytest.mark.parametrize("mode,magnus,pts", [("none", False, 37), ("midpoint", False, 38), ("simpson", False, 41),
                                             ("simpson", False, 40), ("simpson", True, 41)])
def test_parallel_expm_product_equals_sequential(mode, magnus, pts):
    """The block-parallel truth oracle: each block forms its own slices from its sample window
    (shared boundary samples included), so the product equals the sequential one up to the
    association of the matrix products."""
    from oracle.expm_stepwise import expm_stepwise_parallel
    rng = np.random.default_rng(31)
    h = [gue(rng, 5) for _ in range(3)]
    c = rng.uniform(-1, 1, (2, pts))
    seq = expm_stepwise(h[0], h[1:], c, 0.05, mode, magnus)
    for workers in (1, 3, 7):
        par = expm_stepwise_parallel(h[0], h[1:], c, 0.05, mode, magnus, workers=workers)
        assert rel_err(par, seq) < 1e-14
