"""Parity of the HIP equiprop path with the oracle (FP64 restatement of the reference
algorithm, same truncation M and quadrature) through the C ABI via the Parament class.

Bar: relative Frobenius error <= 1e-5 for complex64 contexts and <= 1e-12 for complex128.
The oracle for a complex64 run is evaluated on the complex64-rounded inputs, in FP64,
with the FP32 iteration table.  Precision labels: "fp32" is a complex64 context in its
default arithmetic (the mixed pipeline for dims above 16), "fp32n" a complex64 context in
NATIVE mode (FP32 throughout, the reference's arithmetic), "fp64" a complex128 context."""
import numpy as np
import pytest
import scipy.linalg

from oracle import equiprop_oracle as eo
from tests.helpers import as_c64_inputs, gue, random_hamiltonians, rel_err

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-5, "fp32n": 1e-5, "fp64": 1e-12}
PRECS = ["fp32", "fp32n", "fp64"]


def run_gpu(prec, h0, hks, c, dt, mode="none", magnus=False, path="auto", cycles=None, mixed=False):
    import parament_amd as parament
    with parament.Parament("fp32" if prec == "fp32n" else prec) as p:
        p.set_kernel_path(path)
        if prec == "fp32n":
            p.set_precision_mode("native")
        if mixed:
            p.set_precision_mode("mixed" if mixed is True else mixed)
        if cycles is not None:
            p.set_iteration_cycles(cycles)
        p.set_hamiltonian(h0, *hks, use_magnus=magnus, quadrature_mode=mode)
        u = p.equiprop(dt, *c)
        return u, p.iteration_cycles


def oracle_for(prec, h0, hks, c, dt, mode="none", magnus=False, cycles=None):
    if prec in ("fp32", "fp32n"):
        h0, hks, c = as_c64_inputs(h0, hks, c)
        prec = "fp32"
    return eo.equiprop(h0, hks, c, dt, mode=mode, use_magnus=magnus, precision=prec, cycles=cycles)


def case(seed, d, amps, pts):
    rng = np.random.default_rng(seed)
    h = [gue(rng, d) for _ in range(amps + 1)]
    c = rng.uniform(-1, 1, (amps, pts))
    return h[0], h[1:], c


# ---- dims <= 16: the wavefront-resident kernel folds the lowest tree levels on chip ----

@pytest.mark.parametrize("prec, d, amps, pts, mode", [
    ("fp32", 8, 1, 8195, "none"),            # one folded level, ragged
    ("fp64", 8, 1, 16389, "none"),           # two folded levels, ragged
    ("fp64", 5, 2, 131075, "none"),          # five folded levels, ragged
    ("fp64", 16, 3, 2 * 16389 + 1, "simpson"),
    ("fp64", 3, 2, 20001, "midpoint"),
    # several slices per tile (d <= 8: two, d <= 4: four) with a partly empty last group
    ("fp32", 4, 1, 4099, "none"),
    ("fp64", 2, 1, 7, "none"),
    ("fp32", 4, 2, 3 * 8192 + 5, "none"),
    ("fp64", 7, 3, 2 * 8192 + 3, "simpson"),
    ("fp32", 8, 2, 1 << 16, "none"),
])
def test_small_dim_folded_tree_levels(prec, d, amps, pts, mode):
    """Same pairing and odd carry as the tree kernels: the folded nodes feed the tree above."""
    h0, hks, c = case(4000 + d + pts, d, amps, pts)
    u, m = run_gpu(prec, h0, hks, c, 1e-3, mode)
    assert rel_err(u, oracle_for(prec, h0, hks, c, 1e-3, mode, cycles=m)) < TOL[prec]


# ---- the reference's own numerics tests (test_numerics.py), run through the HIP path ----

@pytest.mark.parametrize("dim", [2, 16])
def test_fp32_expm_scipy_random(dim):
    h0, h1 = random_hamiltonians(dim)
    u, _ = run_gpu("fp32", h0, [h1], np.zeros((1, 1)), 0.01)
    assert np.sum(np.abs(u - scipy.linalg.expm(-1j * 0.01 * h0))) < np.finfo(np.float32).eps * dim * dim


@pytest.mark.parametrize("dim", [2, 16])
def test_fp64_expm_scipy_random(dim):
    h0, h1 = random_hamiltonians(dim)
    u, _ = run_gpu("fp64", h0, [h1], np.zeros((1, 1)), 0.01)
    assert np.sum(np.abs(u - scipy.linalg.expm(-1j * 0.01 * h0))) < np.finfo(np.float64).eps * dim * dim


@pytest.mark.parametrize("dim", [2, 4])
def test_fp32_expm_debug(dim):
    import parament_amd as parament
    h0, _ = random_hamiltonians(dim)
    u = parament.expm(h0)
    assert np.sum(np.abs(u - scipy.linalg.expm(h0))) < np.finfo(np.float32).eps * dim * dim


@pytest.mark.parametrize("dim", [2, 16])
def test_fp32_multi_fields(dim, golden):
    h0, h1 = random_hamiltonians(dim)
    c = golden[f"multi_d{dim}_c"]
    u, _ = run_gpu("fp32", h0, [h1, h0], c, 0.01)
    assert np.linalg.norm(golden[f"multi_d{dim}_expected"] - u) < 1e-6


def test_docstring_example(golden):
    u, m = run_gpu("fp32", golden["doc_h0"], [golden["doc_h1"]], np.zeros((1, 1)), 1.0)
    assert np.allclose(u, golden["doc_expected"], atol=1e-7)
    assert m == eo.select_iteration_cycles_fp32(2.0, 1.0)  # Hnorm = |diag(1,-1)| + |sigma_x| = 2


# ---- oracle parity across modes, dims, paths ----

MODES = [("none", False), ("midpoint", False), ("simpson", False), ("simpson", True)]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("mode,magnus", MODES)
@pytest.mark.parametrize("d", [2, 3, 8, 16, 17, 24, 32, 33, 47, 63, 64])
def test_parity_fused(prec, mode, magnus, d):
    h0, hks, c = case(100 + d, d, 2, 65)
    u, m = run_gpu(prec, h0, hks, c, 0.02, mode, magnus)
    ref = oracle_for(prec, h0, hks, c, 0.02, mode, magnus)
    assert rel_err(u, ref) < TOL[prec], (m,)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("mode,magnus", MODES)
@pytest.mark.parametrize("d", [2, 8, 33, 72, 100, 129, 160])
def test_parity_stepwise(prec, mode, magnus, d):
    h0, hks, c = case(200 + d, d, 2, 33)
    u, _ = run_gpu(prec, h0, hks, c, 0.02, mode, magnus, path="stepwise")
    ref = oracle_for(prec, h0, hks, c, 0.02, mode, magnus)
    assert rel_err(u, ref) < TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("mode,magnus", MODES)
@pytest.mark.parametrize("pts", [1, 2, 3, 4, 5, 9, 10])
def test_edge_pts(prec, mode, magnus, pts):
    import parament_amd as parament
    h0, hks, c = case(300 + pts, 4, 2, pts)
    n = eo.slice_count(pts, eo._MODE_BY_NAME[mode], magnus)
    if n == 0:
        with pytest.raises(ValueError, match="Invalid value"):
            run_gpu(prec, h0, hks, c, 0.05, mode, magnus)
        return
    u, _ = run_gpu(prec, h0, hks, c, 0.05, mode, magnus)
    assert rel_err(u, oracle_for(prec, h0, hks, c, 0.05, mode, magnus)) < TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("cycles", [3, 5, 9, 21])
def test_manual_cycles(prec, cycles):
    h0, hks, c = case(400 + cycles, 8, 1, 40)
    u, m = run_gpu(prec, h0, hks, c, 0.05, "midpoint", cycles=cycles)
    assert m == cycles
    assert rel_err(u, oracle_for(prec, h0, hks, c, 0.05, "midpoint", cycles=cycles)) < TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("amps,magnus", [(256, False), (20, True)])
def test_maximum_coefficients(prec, amps, magnus):
    """The largest coefficient counts a slice holds: 256 controls, or 20 controls with their
    20 + 190 Magnus commutators."""
    h0, hks, c = case(900 + amps, 4, amps, 9)
    u, _ = run_gpu(prec, h0, hks, c, 1e-3, "simpson", magnus)
    assert rel_err(u, oracle_for(prec, h0, hks, c, 1e-3, "simpson", magnus)) < TOL[prec]


def test_fewer_amplitudes_than_hamiltonians():
    h0, hks, c = case(500, 8, 3, 21)
    u, _ = run_gpu("fp64", h0, hks, c[:1], 0.05, "simpson")
    assert rel_err(u, eo.equiprop(h0, hks, c[:1], 0.05, mode="simpson")) < 1e-12


def test_select_smaller_dt():
    h0, hks, c = case(600, 4, 1, 8)
    with pytest.raises(RuntimeError, match="Timestep too large"):
        run_gpu("fp32", 30 * h0, [30 * hks[0]], c, 1.0)


def test_context_reuse_grows_and_shrinks():
    """Buffers grow with pts and are reused for smaller calls (parament.cpp:418 semantics)."""
    import parament_amd as parament
    h0, hks, c = case(700, 16, 2, 257)
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(h0, *hks, quadrature_mode="simpson")
        for pts in (17, 257, 33, 5):
            u = p.equiprop(0.01, *c[:, :pts])
            assert rel_err(u, eo.equiprop(h0, hks, c[:, :pts], 0.01, mode="simpson")) < 1e-12
        # new Hamiltonian of a different dim on the same context
        h0b, hksb, cb = case(701, 33, 1, 9)
        p.set_hamiltonian(h0b, *hksb)
        assert rel_err(p.equiprop(0.01, *cb), eo.equiprop(h0b, hksb, cb, 0.01)) < 1e-12


# ---- reference-config-1 at full size: d = 8, 4096 steps, no quadrature ----

@pytest.mark.parametrize("prec", PRECS)
def test_full_size_dim8_4096_steps(prec):
    rng = np.random.default_rng(20260001)
    h0, h1 = gue(rng, 8), gue(rng, 8)
    c = rng.uniform(-1, 1, (1, 4096))
    for path in ("auto", "stepwise"):
        u, m = run_gpu(prec, h0, [h1], c, 1e-3, "none", path=path)
        assert rel_err(u, oracle_for(prec, h0, [h1], c, 1e-3, "none")) < TOL[prec], (path, m)


# ---- mixed precision (complex128 interface; bf16-split matrix cores + FP64 first-order terms) ----

@pytest.mark.parametrize("mode,magnus", MODES)
@pytest.mark.parametrize("d", [2, 8, 16, 31, 32, 33, 50, 63, 64])
def test_mixed_parity(mode, magnus, d):
    """Same truncation as the FP32 table, within 1e-5 relative of the FP64 restatement."""
    h0, hks, c = case(800 + d, d, 2, 65)
    u, m = run_gpu("fp64", h0, hks, c, 0.02, mode, magnus, mixed=True)
    tau = 0.04 if mode == "simpson" else 0.02
    n = eo.slice_count(65, eo._MODE_BY_NAME[mode], magnus)
    assert m == eo.select_iteration_cycles_mixed(eo.hamiltonian_norm(h0, hks), tau, n)
    ref = eo.equiprop(h0, hks, c, 0.02, mode=mode, use_magnus=magnus, cycles=m)
    assert rel_err(u, ref) < 1e-5


@pytest.mark.parametrize("amps", [3, 4, 6])
@pytest.mark.parametrize("d", [32, 48, 64])
def test_mixed_parity_control_counts(d, amps):
    """The fp16 split kernel assembles the slice Hamiltonian with every operator load of an
    iteration issued at once for up to three operators (a fourth and later slot reads H0 again,
    weighted by zero) and in a loop beyond; both against the FP64 restatement, with complex
    controls of modulus up to 1 (power form), Simpson quadrature."""
    rng = np.random.default_rng(5300 + 10 * d + amps)
    h = [gue(rng, d) for _ in range(amps + 1)]
    c = rng.uniform(-1, 1, (amps, 257)) * np.exp(1j * rng.uniform(0, 2 * np.pi, (amps, 257)))
    u, m = run_gpu("fp64", h[0], h[1:], c, 0.01, "simpson", mixed=True)
    ref = eo.equiprop(h[0], h[1:], c, 0.01, mode="simpson", cycles=m)
    assert rel_err(u, ref) < 1e-5


@pytest.mark.parametrize("d,n", [(d, n) for d in (8, 64) for n in (255, 256, 257, 259, 1024, 1025, 1026, 1027, 1029)]
                         + [(8, n) for n in (4096, 4097, 4098, 4099, 4101, 4103, 4105)] + [(64, 4099), (64, 4103)])
def test_mixed_fused_tree_levels(d, n):
    """The mixed kernel folds the lowest tree levels (1 from 256 slices, 2 from 1024, 3 from
    4096); ragged slice counts exercise the zero-node padding of the last block."""
    h0, hks, c = case(1100 + n + d, d, 2, n)
    u, m = run_gpu("fp64", h0, hks, c, 0.01, "none", mixed=True)
    ref = eo.equiprop(h0, hks, c, 0.01, mode="none", cycles=m)
    assert rel_err(u, ref) < 1e-5


@pytest.mark.parametrize("prec", ["fp32n", "fp64"])
@pytest.mark.parametrize("d,n,dt", [(40, 2051, 0.02), (48, 1030, 0.02), (33, 1029, 0.12), (48, 4099, 0.01)])
def test_fused_48_class_tree_levels(prec, d, n, dt):
    """Dims 33-48 run the 48-wide fused slice kernels (nine waves of 16 x 16 tiles) and the
    48-wide tree: eight-node block levels while more than 1,024 nodes remain, then pairwise
    levels split over work-groups.  Complex64 NATIVE takes the power form for the small steps;
    complex128, and dt = 0.12 (tau Hnorm = 1.14), the recurrence.  (At tau Hnorm = 2.8 plain
    complex64 arithmetic, numpy's included, lands 4.6e-5 from the oracle, beyond the bar.)"""
    h0, hks, c = case(1150 + n + d, d, 2, n)
    u, m = run_gpu(prec, h0, hks, c, dt, "none")
    ref = oracle_for(prec, h0, hks, c, dt, "none")
    assert rel_err(u, ref) < TOL[prec], (m,)


@pytest.mark.parametrize("d,n", [(16, 40), (64, 300), (64, 1100)])
def test_mixed_large_step_keeps_recurrence(d, n):
    """tau Hnorm > 1 selects the Clenshaw recurrence of the mixed kernel (small steps use the
    power form of the same polynomial); both are checked against the FP64 restatement."""
    h0, hks, c = case(1200 + d + n, d, 2, n)
    hn = eo.hamiltonian_norm(h0, hks)
    dt = 2.5 / hn
    u, m = run_gpu("fp64", h0, hks, c, dt, "none", mixed=True)
    assert m >= 9
    ref = eo.equiprop(h0, hks, c, dt, mode="none", cycles=m)
    assert rel_err(u, ref) < 1e-5


@pytest.mark.parametrize("cycles", [3, 5, 9])
def test_mixed_manual_cycles(cycles):
    h0, hks, c = case(900 + cycles, 16, 1, 40)
    u, m = run_gpu("fp64", h0, hks, c, 0.05, "midpoint", cycles=cycles, mixed=True)
    assert m == cycles
    assert rel_err(u, eo.equiprop(h0, hks, c, 0.05, mode="midpoint", cycles=cycles)) < 1e-5


def test_mixed_long_product_dim32():
    """4096 slices of a dim-32 system: error of the mixed path vs the FP64 GPU path."""
    rng = np.random.default_rng(20260004)
    h = [gue(rng, 32) for _ in range(5)]
    c = rng.uniform(-1, 1, (4, 4096))
    u64, _ = run_gpu("fp64", h[0], h[1:], c, 2e-3, "none")
    um, _ = run_gpu("fp64", h[0], h[1:], c, 2e-3, "none", mixed=True)
    assert rel_err(um, u64) < 1e-5


# ---- the headline workload at full size against the oracle: dim-64 Heisenberg chain, 3 controls,
# 65,536 steps, Simpson (32,767 slices).  Each oracle evaluation takes 15-30 s on the host. ----

def _headline(rounded):
    from parament_amd import workloads
    h0, hk, c, dt, quad = workloads.heisenberg6()
    if rounded:
        h0, hk, c = as_c64_inputs(h0, hk, c)
    return h0, hk, c, dt, quad


def test_full_size_heisenberg6_fp64_vs_oracle():
    """complex128 context (FP64 throughout, M = 11 from the FP64 table) against the oracle at the
    same truncation: the complex128 bar, and unitary to 1e-9."""
    h0, hk, c, dt, quad = _headline(False)
    u, m = run_gpu("fp64", h0, hk, c, dt, quad)
    assert m == 11
    assert rel_err(u, eo.equiprop(h0, hk, c, dt, mode=quad, cycles=m)) < 1e-12
    assert np.linalg.norm(u.conj().T @ u - np.eye(64)) < 1e-9


def test_full_size_heisenberg6_mixed_vs_oracle():
    """complex128 context in the mixed mode (M = 7: the FP32 table, which already bounds the
    accumulated truncation error) against the oracle at M = 7: within the 1e-5 gate, and unitary
    to what a deviation of 1e-5 of |U|_F allows, 2 * 1e-5 * sqrt(64)."""
    h0, hk, c, dt, quad = _headline(False)
    u, m = run_gpu("fp64", h0, hk, c, dt, quad, mixed=True)
    assert m == 7
    assert rel_err(u, eo.equiprop(h0, hk, c, dt, mode=quad, cycles=m)) < 1e-5
    assert np.linalg.norm(u.conj().T @ u - np.eye(64)) < 2 * 1e-5 * np.sqrt(64)


def test_full_size_heisenberg6_mixed_larger_controls():
    """The mixed path's error grows with the slices' actual norm, which M and the power form do not
    see (Hnorm leaves out the control amplitudes, as in the reference): with the headline controls
    scaled by 4 (max ||tau X_j||_2 = 0.71 against 0.22) it measured 1.2e-5 against FP64 at the
    same M.  Held below 2e-5, finite, and above the headline's error (so the scaled controls did
    reach the kernel)."""
    h0, hk, c, dt, quad = _headline(False)
    u4, m = run_gpu("fp64", h0, hk, 4.0 * c, dt, quad, mixed=True)
    ref4, _ = run_gpu("fp64", h0, hk, 4.0 * c, dt, quad, cycles=m)
    u1, _ = run_gpu("fp64", h0, hk, c, dt, quad, mixed=True)
    ref1, _ = run_gpu("fp64", h0, hk, c, dt, quad, cycles=m)
    e4, e1 = rel_err(u4, ref4), rel_err(u1, ref1)
    assert np.isfinite(u4).all()
    assert e1 < e4 < 2e-5, (e1, e4)


def test_full_size_heisenberg6_complex64_vs_oracle():
    """complex64 context (the reference's fp32 interface, M = 7 from the FP32 table) against the
    oracle on the complex64-rounded inputs at M = 7.  The default arithmetic (FP64 first-order
    terms and upper tree levels, FP32-accurate split products) meets the complex64 bar of 1e-5.
    FP32 throughout (NATIVE, the reference's arithmetic) measures 7.3e-5 here: roundings that
    are the same in every slice (of s, the a_k and the constant entries of the first-order term)
    add up linearly over the 32,767 slices; the test holds it to 1e-4."""
    h0, hk, c, dt, quad = _headline(True)
    ref = eo.equiprop(h0, hk, c, dt, mode=quad, cycles=7)
    u, m = run_gpu("fp32", h0, hk, c, dt, quad)
    assert m == 7
    assert rel_err(u, ref) < 1e-5
    un, mn = run_gpu("fp32n", h0, hk, c, dt, quad)
    assert mn == 7
    assert rel_err(un, ref) < 1e-4


def test_full_size_ising8_split_products():
    """The precision-variant workload at full size (dim-256 transverse-field Ising chain, 8 local
    sigma_x controls, 16,384 steps, no quadrature) on the batched GEMMs: the default mixed mode
    (the power form on scaled fp16 two-way split products here) and FP32 products, with FP64
    first-order terms and upper tree levels, stay within 1e-6 of the FP64 path; the FP64
    propagator is unitary to 1e-10."""
    import parament_amd as parament
    from parament_amd import workloads
    h0, hk, c, dt, quad = workloads.ising8()
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(h0, *hk, quadrature_mode=quad)
        u64 = p.equiprop(dt, *c)
        errs = {}
        for mode in ("mixed", "mixed_fp32", "mixed_bf16x2"):
            p.set_precision_mode(mode)
            errs[mode] = rel_err(p.equiprop(dt, *c), u64)
    assert np.linalg.norm(u64.conj().T @ u64 - np.eye(256)) < 1e-10
    assert errs["mixed"] < 1e-6 and errs["mixed_fp32"] < 1e-6, errs
    # bf16 hi+lo: each product carries a relative error of up to 3 * 2^-18 (the dropped lo x lo
    # product and the two split residuals); it forms only the O((tau |H|)^2) part of each slice
    # and the three lowest tree levels, and its errors vary from slice to slice, so over 16,384
    # slices they stay below the per-product bound (measured 5.0e-6); above 1e-7 it proves the
    # low-precision products ran
    assert 1e-7 < errs["mixed_bf16x2"] < BF16X2_TOL, errs


# ---- dims above the on-chip kernels: batched matrix-core GEMMs ----

@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("d,n", [(65, 9), (96, 17), (130, 6), (256, 3)])
def test_stepwise_large_dims(prec, d, n):
    """Dims above 64 run the step-wise path on the batched GEMM kernel (64x64 tiles, ragged
    edges, odd tree carries)."""
    h0, hks, c = case(1300 + d, d, 2, n)
    u, m = run_gpu(prec, h0, hks, c, 0.02, "none")
    assert rel_err(u, oracle_for(prec, h0, hks, c, 0.02, "none")) < TOL[prec], m


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("d,n", [(33, 9), (40, 4), (48, 17), (105, 7), (112, 3)])
def test_stepwise_whole_product_dims(prec, d, n):
    """Complex64 products at dims 33-48 and 97-112 run the whole-product GEMM kernel (one
    work-group per slice, 16-row bands); complex128 the tiled kernel."""
    h0, hks, c = case(1600 + d, d, 2, n)
    u, m = run_gpu(prec, h0, hks, c, 0.02, "none", path="stepwise")
    assert rel_err(u, oracle_for(prec, h0, hks, c, 0.02, "none")) < TOL[prec], m


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("cycles", [3, 5, 7, 9, 11])
def test_stepwise_manual_cycles(prec, cycles):
    """The batched recurrence takes its start values b_M, b_(M-1) through C_in in its first two
    steps; with M = 3 the second of them is already the last.  Complex64 contexts take the power
    form from M = 5 on (tau Hnorm <= 1 here): Horner's rule in S = Y^2 over two or three
    alternating buffers, the last step in place over X."""
    h0, hks, c = case(1450 + cycles, 70, 2, 11)
    u, m = run_gpu(prec, h0, hks, c, 0.01, "none", cycles=cycles)
    assert m == cycles
    assert rel_err(u, oracle_for(prec, h0, hks, c, 0.01, "none", cycles=cycles)) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp32", "fp32n"])
def test_stepwise_power_form_tiny_step(prec):
    """A manual M far above the table's at a tiny tau Hnorm: x^M would underflow the power form's
    coefficients, so the step-wise path keeps the recurrence and the result stays finite."""
    h0, hks, c = case(1460, 70, 2, 11)
    u, m = run_gpu(prec, h0, hks, c, 1e-12, "none", cycles=29)
    assert m == 29 and np.all(np.isfinite(u))
    assert rel_err(u, oracle_for(prec, h0, hks, c, 1e-12, "none", cycles=29)) < TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("mode,magnus", MODES)
def test_stepwise_forced_small_dim(prec, mode, magnus):
    h0, hks, c = case(1400, 24, 2, 33)
    u, m = run_gpu(prec, h0, hks, c, 0.03, mode, magnus, path="stepwise")
    assert rel_err(u, oracle_for(prec, h0, hks, c, 0.03, mode, magnus)) < TOL[prec], m


# bf16 hi+lo products: x = hi + lo + r with |r| <= 2^-18 |x|, and the product drops lo x lo, so
# each product is accurate to ~3 * 2^-18 = 1.1e-5 relative; the bound below allows twice that
BF16X2_TOL = 2.5e-5
# what separates a low-precision product mode from FP64 arithmetic: its errors stay above these
FLOOR = {"mixed": 1e-13, "mixed_fp32": 1e-13, "mixed_bf16x2": 1e-9}


@pytest.mark.parametrize("variant,tol", [("mixed", 1e-5), ("mixed_fp32", 1e-5), ("mixed_bf16x2", BF16X2_TOL)])
@pytest.mark.parametrize("d,n,mode", [(65, 40, "simpson"), (100, 33, "none"), (128, 20, "midpoint"), (20, 70, "none")])
def test_mixed_stepwise_modes(variant, tol, d, n, mode):
    """Mixed precision on batched GEMMs (dims above 64, and every dim for the FP32 and bf16
    hi+lo products): Frobenius error against the FP64 restatement with the same truncation, and
    above the floor that shows the low-precision products ran (at d = 20 the FP32 and bf16 hi+lo
    modes take the batched GEMMs, not the FP64 fused kernel)."""
    h0, hks, c = case(1500 + d + n, d, 2, n)
    u, m = run_gpu("fp64", h0, hks, c, 0.01, mode, mixed=variant)
    ref = eo.equiprop(h0, hks, c, 0.01, mode=mode, cycles=m)
    assert FLOOR[variant] < rel_err(u, ref) < tol


@pytest.mark.parametrize("variant,tol", [("mixed", 1e-5), ("mixed_fp32", 1e-5), ("mixed_bf16x2", BF16X2_TOL)])
def test_mixed_stepwise_magnus(variant, tol):
    """The three product modes on batched GEMMs with Magnus slices (commutator operators built on
    the GEMMs, their coefficients formed per slice)."""
    h0, hks, c = case(1600, 70, 2, 33)
    u, m = run_gpu("fp64", h0, hks, c, 0.01, "simpson", True, mixed=variant)
    ref = eo.equiprop(h0, hks, c, 0.01, mode="simpson", use_magnus=True, cycles=m)
    assert FLOOR[variant] < rel_err(u, ref) < tol


def test_mixed_split_scheme_variable():
    """PARAMENT_MIXED_SPLIT=bf16x3 (read once per process, so set in a child process) keeps the bf16
    three-way split in the fused kernel (dim 40) and in the step-wise power form (dim 72), where the
    default runs scaled fp16 two-way splits: both forms within 1e-5 of the oracle, and apart from
    each other by more than FP64 rounding (the variable took effect)."""
    import json
    import os
    import subprocess
    import sys
    code = (
        "import json, sys\n"
        "from tests.test_gpu_parity import case, run_gpu\n"
        "out = {}\n"
        "for d in (40, 72):\n"
        "    h0, hks, c = case(1800 + d, d, 2, 33)\n"
        "    u, m = run_gpu('fp64', h0, hks, c, 0.01, 'none', mixed=True)\n"
        "    out[str(d)] = [u.real.tolist(), u.imag.tolist(), m]\n"
        "print(json.dumps(out))\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, PARAMENT_MIXED_SPLIT="bf16x3")
    r = subprocess.run([sys.executable, "-c", code], cwd=root, env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    other = json.loads(r.stdout.strip().splitlines()[-1])
    for d in (40, 72):
        h0, hks, c = case(1800 + d, d, 2, 33)
        u, m = run_gpu("fp64", h0, hks, c, 0.01, "none", mixed=True)
        re, im, m2 = other[str(d)]
        u2 = np.asarray(re) + 1j * np.asarray(im)
        ref = eo.equiprop(h0, hks, c, 0.01, mode="none", cycles=m)
        assert m2 == m
        assert rel_err(u, ref) < 1e-5 and rel_err(u2, ref) < 1e-5, (d, rel_err(u, ref), rel_err(u2, ref))
        assert rel_err(u, u2) > 1e-10, d


@pytest.mark.parametrize("variant", ["mixed", "mixed_fp32", "mixed_bf16x2"])
@pytest.mark.parametrize("d,mode,magnus", [(8, "none", False), (40, "simpson", True), (72, "midpoint", False)])
def test_complex64_context_product_modes(variant, d, mode, magnus):
    """A complex64 context in each mixed mode: the FP64 copy of its operators (Magnus
    commutators formed again in FP64), its controls widened on the device and the result rounded
    back, against the oracle on the complex64-rounded inputs at the FP32 table's M."""
    h0, hks, c = case(1700 + d, d, 2, 41)
    u, m = run_gpu("fp32", h0, hks, c, 0.01, mode, magnus, mixed=variant)
    assert u.dtype == np.complex64
    ref = oracle_for("fp32", h0, hks, c, 0.01, mode, magnus, cycles=m)
    assert m == eo.select_iteration_cycles_fp32(eo.hamiltonian_norm(*as_c64_inputs(h0, hks, c)[:2]),
                                                0.02 if mode == "simpson" else 0.01)
    assert rel_err(u, ref) < (BF16X2_TOL if variant == "mixed_bf16x2" else 1e-5)


def test_caller_streams_and_concurrent_contexts():
    """Two contexts (a fused and a step-wise dim) on two caller streams at once: the stream a
    context reports is the one it was given, each result equals the oracle once its own stream
    is synchronised, and passing back the handle read first restores the context's own stream."""
    import torch
    import parament_amd as parament
    dev = torch.device("cuda:0")
    cases = [case(800 + d, d, 2, 40) for d in (16, 40)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]   # default priority
    ctxs = []
    try:
        for (h0, hks, c), s in zip(cases, streams):
            p = parament.Parament("fp64")
            ctxs.append(p)
            p.set_hamiltonian(h0, *hks)
            own = p.stream
            assert own
            p.set_stream(s.cuda_stream)
            assert p.stream == s.cuda_stream
            p.own = own
        bufs = []
        for p, (h0, hks, c), s in zip(ctxs, cases, streams):
            d = h0.shape[0]
            with torch.cuda.stream(s):
                carr = torch.tensor(c, dtype=torch.complex128, device=dev)
                out = torch.empty(d * d, dtype=torch.complex128, device=dev)
            p.equiprop_device(0.02, carr.data_ptr(), c.shape[1], c.shape[0], out.data_ptr())
            bufs.append((carr, out))
        for s in streams:
            s.synchronize()
        for (carr, out), (h0, hks, c) in zip(bufs, cases):
            d = h0.shape[0]
            u = out.cpu().numpy().reshape(d, d)
            assert rel_err(u, oracle_for("fp64", h0, hks, c, 0.02)) < TOL["fp64"]
        for p, (h0, hks, c) in zip(ctxs, cases):
            p.set_stream(p.own)
            assert p.stream == p.own
            assert rel_err(p.equiprop(0.02, *c), oracle_for("fp64", h0, hks, c, 0.02)) < TOL["fp64"]
    finally:
        torch.cuda.synchronize()
        for p in ctxs:
            p.destroy()


@pytest.mark.parametrize("prec,d,mode,quad", [("fp32", 8, None, "none"), ("fp64", 8, None, "none"),
                                             ("fp64", 64, "mixed", "simpson"), ("fp32", 40, None, "simpson"),
                                             ("fp64", 72, None, "midpoint"), ("fp32n", 16, None, "none")])
def test_repeated_calls_replay_identically(prec, d, mode, quad):
    """Repeated identical calls are captured as a graph from the third call on and replayed: every
    replay equals the direct call bit for bit (with other work and synchronisations between the
    replays); a replay reads the controls' current contents; a larger call in between (which grows
    the working buffers) does not leave a replay pointing at the old ones."""
    import torch
    import parament_amd as parament
    h0, hks, c = case(2100 + d, d, 2, 257)
    ctype = torch.complex64 if prec.startswith("fp32") else torch.complex128
    with parament.Parament("fp32" if prec == "fp32n" else prec) as p:
        if prec == "fp32n":
            p.set_precision_mode("native")
        if mode:
            p.set_precision_mode(mode)
        p.set_hamiltonian(h0, *hks, quadrature_mode=quad)
        direct = p.equiprop(0.01, *c)
        m = p.iteration_cycles
        tol = 1e-5 if mode == "mixed" else TOL[prec]   # the mixed mode's bar, at its own M
        assert rel_err(direct, oracle_for(prec, h0, hks, c, 0.01, quad, cycles=m)) < tol
        for rep in range(5):
            u = p.equiprop(0.01, *c)
            assert np.array_equal(u, direct), rep
            torch.cuda.synchronize()
            _ = (torch.ones(1000, device="cuda:0") * rep).sum().item()   # unrelated device work in between
        # device pointers: in-place changes of the controls between replays are seen
        carr = torch.tensor(c, dtype=ctype, device="cuda:0")
        out = torch.empty(d * d, dtype=ctype, device="cuda:0")
        for rep in range(4):
            cc = c * (1.0 + 0.1 * rep)
            carr.copy_(torch.tensor(cc, dtype=ctype))
            p.equiprop_device(0.01, carr.data_ptr(), c.shape[1], c.shape[0], out.data_ptr())
            p.synchronize()
            u = out.cpu().numpy().reshape(d, d)
            assert rel_err(u, oracle_for(prec, h0, hks, cc, 0.01, quad, cycles=m)) < tol, rep
        # a larger call grows the buffers; the original call afterwards still matches the direct result
        big = np.tile(c, (1, 4))
        ub = p.equiprop(0.01, *big)
        assert rel_err(ub, oracle_for(prec, h0, hks, big, 0.01, quad, cycles=p.iteration_cycles)) < tol
        for rep in range(3):
            assert np.array_equal(p.equiprop(0.01, *c), direct), rep
