"""Adjoint gate-fidelity gradient on the GPU (Parament.gate_fidelity_gradient) against the
block-matrix-expm oracle (oracle/grape_oracle.py) and against central finite differences of
the library's own FP64 propagator."""
import numpy as np
import pytest

from oracle import grape_oracle as go
from tests.helpers import gue

pytestmark = pytest.mark.gpu


def gpu_grad(h0, hks, c, dt, mode, target, amps=None):
    import parament_amd as parament
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(h0, *hks, quadrature_mode=mode)
        return p.gate_fidelity_gradient(dt, *c, target=target)


def gpu_fidelity(p, c, dt, target):
    u = p.equiprop(dt, *c)
    g = np.trace(target.conj().T @ u)
    return abs(g) ** 2 / target.shape[0] ** 2


@pytest.mark.parametrize("mode", ["none", "midpoint", "simpson"])
@pytest.mark.parametrize("d", [2, 5, 16, 23, 32])
def test_gradient_matches_block_expm_oracle(mode, d):
    from parament_amd.workloads import haar_unitary
    rng = np.random.default_rng(1000 + d)
    h = [gue(rng, d) for _ in range(3)]
    c = rng.uniform(-1, 1, (2, 21))
    t = haar_unitary(rng, d)
    f, grad = gpu_grad(h[0], h[1:], c, 0.05, mode, t)
    f_ref, grad_ref = go.fidelity_gradient(h[0], h[1:], c, 0.05, mode, t)
    assert abs(f - f_ref) < 1e-12
    assert np.abs(grad - grad_ref).max() < 1e-10 * max(1.0, np.abs(grad_ref).max())


def test_gradient_matches_central_differences_dim32_4096_steps():
    """The GRAPE workload at full size: dim 32, 4 GUE controls, 4096 steps, dt 2e-3, Haar
    target; central differences (h = 1e-6) of the FP64 equiprop at 12 random entries."""
    import parament_amd as parament
    from parament_amd import workloads
    h0, hk, c, dt, quad, target = workloads.gue32_grape()
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(h0, *hk, quadrature_mode=quad)
        f, grad = p.gate_fidelity_gradient(dt, *c, target=target)
        assert abs(f - gpu_fidelity(p, c, dt, target)) < 1e-12
        rng = np.random.default_rng(7)
        hstep = 1e-6
        for _ in range(12):
            k, j = int(rng.integers(4)), int(rng.integers(c.shape[1]))
            cp, cm = c.copy(), c.copy()
            cp[k, j] += hstep
            cm[k, j] -= hstep
            fd = (gpu_fidelity(p, cp, dt, target) - gpu_fidelity(p, cm, dt, target)) / (2 * hstep)
            assert abs(fd - grad[k, j]) < 1e-6 * max(1.0, abs(grad).max()), (k, j, fd, grad[k, j])


def test_gradient_fewer_amplitudes():
    from parament_amd.workloads import haar_unitary
    rng = np.random.default_rng(5)
    h = [gue(rng, 4) for _ in range(4)]
    c = rng.uniform(-1, 1, (2, 11))
    t = haar_unitary(rng, 4)
    f, grad = gpu_grad(h[0], h[1:], c, 0.05, "none", t)
    assert grad.shape == (2, 11)
    f_ref, grad_ref = go.fidelity_gradient(h[0], h[1:], c, 0.05, "none", t)
    assert np.abs(grad - grad_ref).max() < 1e-10


def test_gradient_argument_errors():
    import parament_amd as parament
    t = np.eye(2)
    with parament.Parament("fp32") as p:
        p.set_hamiltonian(np.diag([1.0, -1.0]), np.array([[0, 1], [1, 0]]))
        with pytest.raises(ValueError):
            p.gate_fidelity_gradient(0.1, np.zeros(4), target=t)
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(np.diag([1.0, -1.0]), np.array([[0, 1], [1, 0]]), use_magnus=True, quadrature_mode="simpson")
        with pytest.raises(ValueError, match="Invalid value"):
            p.gate_fidelity_gradient(0.1, np.zeros(5), target=t)
        p.set_hamiltonian(np.diag([1.0, -1.0]), np.array([[0, 1], [1, 0]]))
        with pytest.raises(ValueError):
            p.gate_fidelity_gradient(0.1, np.zeros(4) + 1j, target=t)
