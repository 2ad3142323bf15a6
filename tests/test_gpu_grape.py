"""Adjoint gate-fidelity gradient on the GPU (Parament.gate_fidelity_gradient) against the
block-matrix-expm oracle (oracle/grape_oracle.py) and against central finite differences of
the library's own FP64 propagator; checkpointed and unsegmented layouts must agree bit for bit."""
import numpy as np
import pytest

from oracle import grape_oracle as go
from tests.helpers import gue

pytestmark = pytest.mark.gpu


def _problem(seed, d, amps, pts):
    from parament_amd.workloads import haar_unitary
    rng = np.random.default_rng(seed)
    h = [gue(rng, d) for _ in range(amps + 1)]
    c = rng.uniform(-1, 1, (amps, pts))
    return h[0], h[1:], c, haar_unitary(rng, d)


def gpu_grad(prec, h0, hks, c, dt, mode, target, segment=0, chunk=0):
    import parament_amd as parament
    with parament.Parament(prec) as p:
        p.set_hamiltonian(h0, *hks, quadrature_mode=mode)
        if segment:
            p.set_gradient_checkpointing(segment, chunk)
        return p.gate_fidelity_gradient(dt, *c, target=target)


def gpu_fidelity(p, c, dt, target):
    u = p.equiprop(dt, *c)
    g = np.trace(target.conj().T @ u)
    return abs(g) ** 2 / target.shape[0] ** 2


@pytest.mark.parametrize("mode", ["none", "midpoint", "simpson"])
@pytest.mark.parametrize("d", [2, 5, 16, 23, 32])
def test_gradient_matches_block_expm_oracle(mode, d):
    h0, hk, c, t = _problem(1000 + d, d, 2, 21)
    f, grad = gpu_grad("fp64", h0, hk, c, 0.05, mode, t)
    f_ref, grad_ref = go.fidelity_gradient(h0, hk, c, 0.05, mode, t)
    assert abs(f - f_ref) < 1e-12
    assert np.abs(grad - grad_ref).max() < 1e-10 * max(1.0, np.abs(grad_ref).max())


@pytest.mark.parametrize("d", [8, 40])
def test_gradient_many_controls(d):
    """43 controls: the traces run eight controls per pass over each slice's derivative, so
    five full groups and a partial last one are checked against the oracle."""
    h0, hk, c, t = _problem(1700 + d, d, 43, 11)
    hk = [0.2 * m for m in hk]
    f, grad = gpu_grad("fp64", h0, hk, c, 0.02, "none", t)
    f_ref, grad_ref = go.fidelity_gradient(h0, hk, c, 0.02, "none", t)
    assert abs(f - f_ref) < 1e-12
    assert np.abs(grad - grad_ref).max() < 1e-10 * max(1.0, np.abs(grad_ref).max())


@pytest.mark.parametrize("d", [5, 32])
@pytest.mark.parametrize("dt", [0.02, 0.3])
def test_gradient_both_derivative_forms_match_oracle(d, dt):
    """tau Hnorm below 1 (derivative of the power form) and above 1 (forward-mode Clenshaw
    derivative): both are the exact derivative of the propagator."""
    h0, hk, c, t = _problem(1500 + d, d, 3, 17)
    hnorm = sum(np.abs(m).sum(axis=1).max() for m in (h0, *hk))
    assert (dt * hnorm < 1) == (dt == 0.02)
    f, grad = gpu_grad("fp64", h0, hk, c, dt, "none", t)
    f_ref, grad_ref = go.fidelity_gradient(h0, hk, c, dt, "none", t)
    assert abs(f - f_ref) < 1e-12
    assert np.abs(grad - grad_ref).max() < 1e-10 * max(1.0, np.abs(grad_ref).max())


@pytest.mark.parametrize("mode", ["none", "midpoint", "simpson"])
@pytest.mark.parametrize("d", [3, 16, 32])
def test_complex64_gradient_matches_block_expm_oracle(mode, d):
    h0, hk, c, t = _problem(2000 + d, d, 3, 41)
    f, grad = gpu_grad("fp32", h0, hk, c, 0.05, mode, t)
    assert grad.dtype == np.float32
    f_ref, grad_ref = go.fidelity_gradient(h0, hk, c, 0.05, mode, t)
    assert abs(f - f_ref) < 1e-5 * max(f_ref, 1e-3)
    err = np.abs(grad - grad_ref).max() / np.abs(grad_ref).max()
    assert err < 1e-4, err


@pytest.mark.parametrize("mode", ["none", "midpoint", "simpson"])
@pytest.mark.parametrize("d", [33, 48, 70])
def test_stepwise_gradient_matches_block_expm_oracle(mode, d):
    """Dims above 32 take the batched-GEMM gradient (whole tree kept, Clenshaw derivative in the
    direction of each slice's costate, one trace per control)."""
    h0, hk, c, t = _problem(4000 + d, d, 2, 21)
    f, grad = gpu_grad("fp64", h0, hk, c, 0.05, mode, t)
    f_ref, grad_ref = go.fidelity_gradient(h0, hk, c, 0.05, mode, t)
    assert abs(f - f_ref) < 1e-12
    assert np.abs(grad - grad_ref).max() < 1e-10 * max(1.0, np.abs(grad_ref).max())


@pytest.mark.parametrize("d,dt", [(40, 0.02), (40, 0.4), (65, 0.1)])
def test_complex64_stepwise_gradient_matches_block_expm_oracle(d, dt):
    h0, hk, c, t = _problem(4100 + d, d, 3, 41)
    f, grad = gpu_grad("fp32", h0, hk, c, dt, "simpson", t)
    assert grad.dtype == np.float32
    f_ref, grad_ref = go.fidelity_gradient(h0, hk, c, dt, "simpson", t)
    assert abs(f - f_ref) < 1e-5 * max(f_ref, 1e-3)
    err = np.abs(grad - grad_ref).max() / np.abs(grad_ref).max()
    assert err < 1e-4, err


def test_complex64_stepwise_gradient_target_scale():
    """The default complex64 step-wise gradient runs its products on fp16 splits and its costates
    from t^H / ||t||_F: a target scaled by 3 gives 9 times the gradient, a zero target a zero
    gradient, and the FP32-product mode (native) the same gradient."""
    import parament_amd as parament
    h0, hk, c, t = _problem(4150, 48, 2, 33)
    f_ref, grad_ref = go.fidelity_gradient(h0, hk, c, 0.05, "none", t)
    res = {}
    for mode in ("mixed", "native"):
        with parament.Parament("fp32") as p:
            p.set_hamiltonian(h0, *hk)
            p.set_precision_mode(mode)
            assert p.precision_mode == mode
            res[mode] = [p.gate_fidelity_gradient(0.05, *c, target=x) for x in (t, 3.0 * t, np.zeros_like(t))]
    scale = np.abs(grad_ref).max()
    for mode, ((f1, g1), (f3, g3), (f0, g0)) in res.items():
        assert abs(f1 - f_ref) < 1e-5 * max(f_ref, 1e-3), mode
        assert np.abs(g1 - grad_ref).max() < 1e-4 * scale, mode
        assert abs(f3 - 9 * f1) < 1e-5 * 9 * max(f1, 1e-3), mode
        assert np.abs(g3 - 9 * g1).max() < 1e-4 * 9 * scale, mode
        assert f0 == 0.0 and np.all(g0 == 0.0), mode
    assert np.abs(res["mixed"][0][1] - res["native"][0][1]).max() < 1e-4 * scale


@pytest.mark.parametrize("pts,mode", [(1, "none"), (2, "none"), (3, "simpson"), (5, "midpoint"), (33, "none")])
def test_stepwise_gradient_few_slices(pts, mode):
    """One slice (no tree), two, and odd counts that carry a node up a level."""
    h0, hk, c, t = _problem(4200 + pts, 36, 2, pts)
    f, grad = gpu_grad("fp64", h0, hk, c, 0.07, mode, t)
    f_ref, grad_ref = go.fidelity_gradient(h0, hk, c, 0.07, mode, t)
    assert abs(f - f_ref) < 1e-12
    assert np.abs(grad - grad_ref).max() < 1e-10 * max(1.0, np.abs(grad_ref).max())


def _fd_check(p, c, dt, t, grad, entries, h=1e-6, tol=1e-8):
    for k, j in entries:
        cp, cm = c.copy(), c.copy()
        cp[k, j] += h
        cm[k, j] -= h
        fd = (gpu_fidelity(p, cp, dt, t) - gpu_fidelity(p, cm, dt, t)) / (2 * h)
        assert abs(fd - grad[k, j]) < tol * max(1.0, np.abs(grad).max()), (k, j, fd, grad[k, j])


@pytest.mark.parametrize("d", [5, 40])
def test_magnus_gradient_matches_central_differences(d):
    """Magnus slices (Simpson means plus commutator terms whose coefficients are differences and
    products of samples) on the fused (d = 5) and batched-GEMM (d = 40) paths: central
    differences of the library's own FP64 Magnus propagator, at even and odd samples, the first
    and the last, of every control."""
    import parament_amd as parament
    h0, hk, c, t = _problem(4500 + d, d, 3, 21)
    dt = 0.05
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(h0, *hk, use_magnus=True, quadrature_mode="simpson")
        f, grad = p.gate_fidelity_gradient(dt, *c, target=t)
        assert abs(f - gpu_fidelity(p, c, dt, t)) < 1e-13
        _fd_check(p, c, dt, t, grad, [(0, 0), (1, 1), (2, 2), (0, 7), (1, 10), (2, 13), (0, 19), (1, 20), (2, 20)])


@pytest.mark.parametrize("d,dt", [(12, 0.05), (40, 0.05), (40, 0.01)])
def test_complex64_magnus_gradient_matches_fp64(d, dt):
    """Magnus gradients of a complex64 context against the complex128 ones: fused (d = 12), and
    on the batched GEMMs with the recurrence (tau Hnorm = 3.0) and with the power form on fp16
    split products (tau Hnorm = 0.61, M = 7)."""
    import parament_amd as parament
    h0, hk, c, t = _problem(4600, d, 3, 33)
    out = {}
    for prec in ("fp64", "fp32"):
        with parament.Parament(prec) as p:
            p.set_hamiltonian(h0, *hk, use_magnus=True, quadrature_mode="simpson")
            out[prec] = p.gate_fidelity_gradient(dt, *c, target=t)
    (f64, g64), (f32, g32) = out["fp64"], out["fp32"]
    assert abs(f32 - f64) < 1e-5 * max(f64, 1e-3)
    assert np.abs(g32 - g64).max() < 1e-4 * np.abs(g64).max()


@pytest.mark.parametrize("cycles", [3, 5, 7, 9])
def test_stepwise_gradient_manual_cycles(cycles):
    """A manual Chebyshev order (M = 3: the derivative's first two steps are all of it; from M = 5
    on the power form's Horner derivative, whose levels below J - 1 rotate through a spare buffer
    from M = 7 on) against central differences of the library's own FP64 propagator with the
    same order."""
    import parament_amd as parament
    h0, hk, c, t = _problem(4400 + cycles, 40, 2, 15)
    dt, h = 0.01, 1e-6
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(h0, *hk)
        p.set_iteration_cycles(cycles)
        f, grad = p.gate_fidelity_gradient(dt, *c, target=t)
        assert p.iteration_cycles == cycles
        assert abs(f - gpu_fidelity(p, c, dt, t)) < 1e-13
        for k, j in [(0, 0), (1, 7), (0, 14), (1, 3)]:
            cp, cm = c.copy(), c.copy()
            cp[k, j] += h
            cm[k, j] -= h
            fd = (gpu_fidelity(p, cp, dt, t) - gpu_fidelity(p, cm, dt, t)) / (2 * h)
            assert abs(fd - grad[k, j]) < 1e-8 * max(1.0, np.abs(grad).max()), (k, j, fd, grad[k, j])


@pytest.mark.parametrize("d,cycles,dt", [(8, 3, 0.01), (8, 7, 0.01), (32, 5, 0.01), (16, 9, 0.1), (32, 21, 0.2)])
def test_fused_gradient_manual_cycles(d, cycles, dt):
    """A manual Chebyshev order on the fused kernels (the power-form derivative where tau Hnorm
    <= 1, the forward-mode Clenshaw derivative above) against central differences of the
    library's own FP64 propagator with the same order."""
    import parament_amd as parament
    h0, hk, c, t = _problem(4500 + d + cycles, d, 2, 15)
    h = 1e-6
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(h0, *hk)
        p.set_iteration_cycles(cycles)
        f, grad = p.gate_fidelity_gradient(dt, *c, target=t)
        assert p.iteration_cycles == cycles
        assert abs(f - gpu_fidelity(p, c, dt, t)) < 1e-12
        for k, j in [(0, 0), (1, 7), (0, 14), (1, 3)]:
            cp, cm = c.copy(), c.copy()
            cp[k, j] += h
            cm[k, j] -= h
            fd = (gpu_fidelity(p, cp, dt, t) - gpu_fidelity(p, cm, dt, t)) / (2 * h)
            assert abs(fd - grad[k, j]) < 1e-8 * max(1.0, np.abs(grad).max()), (k, j, fd, grad[k, j])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_stepwise_gradient_chunks_agree(prec):
    """Per-slice arrays processed 7 slices at a time give the whole-batch result; the default call
    also keeps the slices' S and Horner brackets for the derivative, the checkpointed one forms
    them again."""
    h0, hk, c, t = _problem(4300, 40, 2, 61)
    f0, g0 = gpu_grad(prec, h0, hk, c, 0.05, "none", t)
    f1, g1 = gpu_grad(prec, h0, hk, c, 0.05, "none", t, segment=1, chunk=7)
    tol = 1e-13 if prec == "fp64" else 1e-6
    assert abs(f0 - f1) <= tol
    assert np.abs(g0 - g1).max() <= tol * max(1.0, np.abs(g0).max())


def test_gradient_matches_central_differences_dim32_4096_steps():
    """The GRAPE workload at full size: dim 32, 4 GUE controls, 4096 steps, dt 2e-3, Haar
    target; central differences (h = 1e-6) of the FP64 equiprop at 16 random entries."""
    import parament_amd as parament
    from parament_amd import workloads
    h0, hk, c, dt, quad, target = workloads.gue32_grape()
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(h0, *hk, quadrature_mode=quad)
        f, grad = p.gate_fidelity_gradient(dt, *c, target=target)
        assert abs(f - gpu_fidelity(p, c, dt, target)) < 1e-12
        rng = np.random.default_rng(7)
        hstep = 1e-6
        worst = 0.0
        for _ in range(16):
            k, j = int(rng.integers(4)), int(rng.integers(c.shape[1]))
            cp, cm = c.copy(), c.copy()
            cp[k, j] += hstep
            cm[k, j] -= hstep
            fd = (gpu_fidelity(p, cp, dt, target) - gpu_fidelity(p, cm, dt, target)) / (2 * hstep)
            worst = max(worst, abs(fd - grad[k, j]))
        scale = np.abs(grad).max()
        print(f"max |FD - adjoint| = {worst:.3e}, max |grad| = {scale:.3e}, ratio {worst / scale:.3e}")
        assert worst < 1e-6
        assert worst < 1e-3 * scale


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gradient_dim32_4096_steps_window_vs_block_expm(prec):
    """The GRAPE workload at full size against the exact-derivative oracle (block-matrix expm,
    the full 4096-slice prefix and suffix products) on a contiguous window of 512 steps in the
    middle of the sequence, every control: FP64 to 1e-10 and complex64 to 1e-3 of the largest
    component (the Chebyshev truncation at tau Hnorm = 0.03 is far below both)."""
    from parament_amd import workloads
    h0, hk, c, dt, quad, target = workloads.gue32_grape()
    f, grad = gpu_grad(prec, h0, hk, c, dt, quad, target)
    win = range(1792, 2304)
    f_ref, grad_ref = go.fidelity_gradient(h0, hk, c, dt, quad, target, slices=win)
    g, r = grad[:, win.start:win.stop], grad_ref[:, win.start:win.stop]
    err = np.abs(g - r).max() / np.abs(r).max()
    print(f"{prec}: window max error / max |grad| = {err:.3e}, F {f:.9e} vs {f_ref:.9e}")
    if prec == "fp64":
        assert abs(f - f_ref) < 1e-12
        assert err < 1e-10
    else:
        assert abs(f - f_ref) < 1e-4 * f_ref
        assert err < 1e-3


@pytest.mark.parametrize("amps,magnus", [(10, True), (70, False)])
def test_gradient_many_coefficient_operators_small_dim(amps, magnus):
    """More coefficient operators than the fused gradient kernels hold (64) at a small dim: 10
    Magnus controls carry 10 + 10 + 45 = 65 operators, 70 plain controls 70; the batched-GEMM
    gradient takes them, against the exact-derivative oracle (Magnus slices included)."""
    h0, hk, c, t = _problem(4800 + amps, 8, amps, 9)
    hk = [m / amps for m in hk]
    mode = "simpson" if magnus else "none"
    import parament_amd as parament
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(h0, *hk, use_magnus=magnus, quadrature_mode=mode)
        f, grad = p.gate_fidelity_gradient(0.05, *c, target=t)
    f_ref, grad_ref = go.fidelity_gradient(h0, hk, c, 0.05, mode, t, use_magnus=magnus)
    assert abs(f - f_ref) < 1e-12
    assert np.abs(grad - grad_ref).max() < 1e-10 * max(1.0, np.abs(grad_ref).max())


def test_complex64_gradient_at_dim32_4096_steps():
    """complex64 gradient of the same workload against the FP64 adjoint gradient (which the
    tests above pin to the block-expm oracle and to finite differences)."""
    from parament_amd import workloads
    h0, hk, c, dt, quad, target = workloads.gue32_grape()
    f64, g64 = gpu_grad("fp64", h0, hk, c, dt, quad, target)
    f32, g32 = gpu_grad("fp32", h0, hk, c, dt, quad, target)
    err = np.abs(g32 - g64).max() / np.abs(g64).max()
    print(f"complex64 gradient: max error / max |grad| = {err:.3e}, F {f32:.6e} vs {f64:.6e}")
    assert abs(f32 - f64) < 1e-4 * f64
    assert err < 1e-3


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("pts,segment,chunk", [
    (1, 1, 0), (2, 1, 0), (3, 2, 2), (7, 2, 2), (64, 1, 8), (65, 1, 8), (1000, 1, 0), (1000, 4, 8),
    (1000, 2, 2), (4097, 1, 0), (4097, 16, 64), (5000, 64, 2048)])
def test_checkpointed_gradient_is_identical(prec, pts, segment, chunk):
    import parament_amd as parament
    h0, hk, c, t = _problem(3000 + pts, 4, 2, pts)
    out = []
    for seg in (0, segment):
        with parament.Parament(prec) as p:
            p.set_hamiltonian(h0, *hk, quadrature_mode="none")
            if seg:
                p.set_gradient_checkpointing(seg, chunk)
            out.append(p.gate_fidelity_gradient(0.01, *c, target=t) + (p.gradient_workspace,))
    (f0, g0, w0), (f1, g1, w1) = out
    assert f0 == f1
    assert np.array_equal(g0, g1)
    assert w1 <= w0


def test_checkpointing_bounds_memory_dim32():
    """sqrt(N) segments with small chunks: the working set shrinks from O(N) to O(sqrt N + chunk)
    matrices and the gradient does not change."""
    import parament_amd as parament
    from parament_amd import workloads
    h0, hk, c, dt, quad, target = workloads.gue32_grape()
    res = {}
    for seg, chunk in ((0, 0), (1, 0), (1, 256)):
        with parament.Parament("fp64") as p:
            p.set_hamiltonian(h0, *hk, quadrature_mode=quad)
            if seg:
                p.set_gradient_checkpointing(seg, chunk)
            f, g = p.gate_fidelity_gradient(dt, *c, target=target)
            res[(seg, chunk)] = (f, g, p.gradient_workspace)
    full = res[(0, 0)]
    for key in ((1, 0), (1, 256)):
        assert res[key][0] == full[0]
        assert np.array_equal(res[key][1], full[1])
    dd16 = 32 * 32 * 16
    print({k: v[2] / dd16 for k, v in res.items()})
    assert res[(1, 256)][2] < full[2] / 8


@pytest.mark.parametrize("d", [4, 40])
def test_gradient_fewer_amplitudes(d):
    """Two of three control Hamiltonians driven: the third has zero amplitude (fused and
    batched-GEMM paths)."""
    h0, hk, c, t = _problem(5 + d, d, 3, 11)
    c = c[:2]
    f, grad = gpu_grad("fp64", h0, hk, c, 0.05, "none", t)
    assert grad.shape == (2, 11)
    f_ref, grad_ref = go.fidelity_gradient(h0, hk[:2], c, 0.05, "none", t)
    assert np.abs(grad - grad_ref).max() < 1e-10 * max(1.0, np.abs(grad_ref).max())


def test_magnus_gradient_fewer_amplitudes():
    """Magnus with two of three controls driven: pair terms with the undriven control vanish."""
    import parament_amd as parament
    h0, hk, c, t = _problem(4700, 6, 3, 15)
    c = c[:2]
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(h0, *hk, use_magnus=True, quadrature_mode="simpson")
        f, grad = p.gate_fidelity_gradient(0.05, *c, target=t)
        assert grad.shape == (2, 15)
        _fd_check(p, c, 0.05, t, grad, [(0, 0), (1, 3), (0, 8), (1, 14)])


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_gradient_argument_errors(prec):
    import parament_amd as parament
    t = np.eye(2)
    sx = np.array([[0, 1], [1, 0]])
    with parament.Parament(prec) as p:
        with pytest.raises(RuntimeError):
            p.gate_fidelity_gradient(0.1, np.zeros(4), target=t)
        p.set_hamiltonian(np.diag([1.0, -1.0]), sx, use_magnus=True, quadrature_mode="simpson")
        f, g = p.gate_fidelity_gradient(0.1, np.zeros(5), target=t)   # Magnus contexts are supported
        assert np.isfinite(f) and np.all(np.isfinite(g))
        p.set_hamiltonian(np.diag([1.0, -1.0]), sx)
        with pytest.raises(ValueError):
            p.gate_fidelity_gradient(0.1, np.zeros(4) + 1j, target=t)
        with pytest.raises(ValueError):
            p.gate_fidelity_gradient(0.1, np.zeros(4), np.zeros(4), target=t)
        big = np.eye(33)
        p.set_hamiltonian(big, big)
        with pytest.raises(ValueError, match="dim x dim"):
            p.gate_fidelity_gradient(0.1, np.zeros(4), target=t)
        # dims above 32 (batched GEMMs): a control on the identity only moves the global phase
        f, g = p.gate_fidelity_gradient(0.1, np.zeros(4), target=big)
        assert abs(f - 1) < 1e-5 and np.abs(g).max() < 1e-4


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("segment,d", [(0, 16), (1, 16), (0, 40)])
def test_repeated_gradient_calls_replay_identically(prec, segment, d):
    """A call repeated with the same arguments is captured once and then replayed as a graph:
    every repeat equals the first (direct) call bit for bit, and a replay reads the buffers'
    current contents (controls changed in place match a fresh context's direct call)."""
    import torch
    import parament_amd as parament
    h0, hk, c, t = _problem(77, d, 3, 300)
    ct = torch.complex128 if prec == "fp64" else torch.complex64
    rt = torch.float64 if prec == "fp64" else torch.float32
    dev = torch.device("cuda:0")
    carr = torch.tensor(c, dtype=ct, device=dev).contiguous()
    tgt = torch.tensor(t, dtype=ct, device=dev).contiguous()

    def run(p, n):
        out = []
        f = torch.empty(1, dtype=rt, device=dev)
        g = torch.empty(c.shape, dtype=rt, device=dev)
        for _ in range(n):
            p.gate_fidelity_gradient_device(0.02, carr.data_ptr(), c.shape[1], c.shape[0], tgt.data_ptr(),
                                            f.data_ptr(), g.data_ptr())
            torch.cuda.synchronize()
            out.append((f.item(), g.cpu().numpy().copy()))
        return out

    with parament.Parament(prec) as p:
        p.set_hamiltonian(h0, *hk, quadrature_mode="simpson")
        p.set_stream(torch.cuda.current_stream().cuda_stream)
        p.set_gradient_checkpointing(segment, 0)
        first, *rest = run(p, 4)
        for f, g in rest:
            assert f == first[0] and np.array_equal(g, first[1])
        carr.mul_(0.5)
        replayed = run(p, 2)
    with parament.Parament(prec) as q:
        q.set_hamiltonian(h0, *hk, quadrature_mode="simpson")
        q.set_stream(torch.cuda.current_stream().cuda_stream)
        q.set_gradient_checkpointing(segment, 0)
        fresh = run(q, 1)[0]
    for f, g in replayed:
        assert f == fresh[0] and np.array_equal(g, fresh[1])
    assert not np.array_equal(fresh[1], first[1])


def _torch_problem(prec, seed, d, amps, pts):
    import torch
    h0, hk, c, t = _problem(seed, d, amps, pts)
    real_t = torch.float64 if prec == "fp64" else torch.float32
    cplx_t = torch.complex128 if prec == "fp64" else torch.complex64
    dev = torch.device("cuda:0")
    return h0, hk, torch.tensor(c, dtype=real_t, device=dev), torch.tensor(t, dtype=cplx_t, device=dev), c, t


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_autograd_matches_adjoint(prec):
    import torch
    import parament_amd as parament
    from parament_amd.grape import gate_fidelity
    h0, hk, ct, tt, c, t = _torch_problem(prec, 11, 8, 2, 33)
    with parament.Parament(prec) as p:
        p.set_hamiltonian(h0, *hk, quadrature_mode="simpson")
        f_ref, g_ref = p.gate_fidelity_gradient(0.03, *c, target=t)
        x = ct.clone().requires_grad_(True)
        f = gate_fidelity(p, x, 0.03, tt)
        (3.0 * f).backward()
        torch.cuda.synchronize()
        assert abs(f.item() - f_ref) <= 1e-6 * f_ref
        np.testing.assert_allclose(x.grad.cpu().numpy(), 3.0 * g_ref, rtol=1e-5, atol=1e-7 * np.abs(g_ref).max())
        with torch.no_grad():
            f2 = gate_fidelity(p, ct, 0.03, tt)
        assert abs(f2.item() - f_ref) <= (1e-12 if prec == "fp64" else 1e-5) * max(f_ref, 1e-3)


def test_autograd_gradcheck():
    import torch
    import parament_amd as parament
    from parament_amd.grape import gate_fidelity
    h0, hk, ct, tt, c, t = _torch_problem("fp64", 12, 3, 2, 9)
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(h0, *hk, quadrature_mode="midpoint")
        x = ct.clone().requires_grad_(True)
        assert torch.autograd.gradcheck(lambda y: gate_fidelity(p, y, 0.05, tt), (x,), eps=1e-6, atol=1e-8, rtol=1e-5)


def test_autograd_optimisation_raises_fidelity():
    """A few Adam steps of a GRAPE loop driven by the autograd wrapper raise the fidelity."""
    import torch
    import parament_amd as parament
    from parament_amd.grape import gate_fidelity
    h0, hk, ct, tt, c, t = _torch_problem("fp64", 13, 4, 3, 200)
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(h0, *hk, quadrature_mode="none")
        x = (0.1 * ct).clone().requires_grad_(True)
        opt = torch.optim.Adam([x], lr=0.05)
        fids = []
        for _ in range(30):
            opt.zero_grad()
            f = gate_fidelity(p, x, 0.02, tt)
            (-f).backward()
            opt.step()
            fids.append(f.item())
        assert fids[-1] > fids[0] + 0.1
