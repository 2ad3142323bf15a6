"""Ensembles of independent complex64 problems (parament_amd.Ensemble): every instance against
the complex64 restatement of the reference algorithm, the captured-graph execution against the
per-instance loop, and repeated graph replays against each other."""
import numpy as np
import pytest

from oracle import equiprop_oracle as eo
from tests.helpers import as_c64_inputs, gue, rel_err

pytestmark = pytest.mark.gpu


def make_ensemble(seed, dims, amps, pts, dt):
    rng = np.random.default_rng(seed)
    hs = [[gue(rng, d) for _ in range(a + 1)] for d, a in zip(dims, amps)]
    cs = [rng.uniform(-1, 1, (a, p)) for a, p in zip(amps, pts)]
    return hs, cs


DIMS = [3, 8, 16, 17, 31, 32, 33, 48, 64, 65, 80, 100, 5, 40]
AMPS = [1, 2, 3, 1, 4, 2, 1, 3, 2, 1, 2, 1, 4, 3]
PTS = [20, 257, 100, 33, 64, 1, 129, 80, 47, 31, 9, 17, 2049, 300]


@pytest.mark.parametrize("quad", ["none", "simpson"])
def test_ensemble_matches_restatement(quad):
    from parament_amd.ensemble import Ensemble
    pts = [p + 2 for p in PTS] if quad == "simpson" else PTS
    dt = [0.01 + 0.002 * i for i in range(len(DIMS))]
    hs, cs = make_ensemble(7, DIMS, AMPS, pts, dt)
    with Ensemble(DIMS, AMPS, pts, dt, quad) as e:
        e.set_hamiltonians(hs)
        us = e.equiprop(cs)
        for i, (h, c, u) in enumerate(zip(hs, cs, us)):
            h0, hk, c32 = as_c64_inputs(h[0], h[1:], c)
            ref = eo.equiprop(h0, hk, c32, dt[i], mode=quad, precision="fp32")
            assert rel_err(u, ref) < 1e-5, (i, DIMS[i], pts[i])


@pytest.mark.parametrize("dims,amps,pts,quad", [
    ([7], [2], [50], "midpoint"),                                   # one instance
    ([20, 24, 30, 18], [1, 2, 1, 3], [64, 33, 100, 5], "none"),     # one class only
    ([90], [1], [41], "simpson"),                                   # one batched-GEMM instance
    ([70, 128, 66], [2, 1, 3], [30, 20, 65], "midpoint"),           # batched-GEMM instances only
])
def test_ensemble_shapes(dims, amps, pts, quad):
    """Ensembles of one instance, of one class, and of batched-GEMM instances only, as the graph
    and as the loop, then again through the same object after new Hamiltonians (the plan and the
    tuning are rebuilt)."""
    from parament_amd.ensemble import Ensemble
    dt = [0.02] * len(dims)
    with Ensemble(dims, amps, pts, dt, quad) as e:
        for seed in (31, 32):
            hs, cs = make_ensemble(seed, dims, amps, pts, dt)
            e.set_hamiltonians(hs)
            for mode in ("graph", "loop"):
                e.set_execution(mode)
                us = e.equiprop(cs)
                for i, (h, c, u) in enumerate(zip(hs, cs, us)):
                    h0, hk, c32 = as_c64_inputs(h[0], h[1:], c)
                    ref = eo.equiprop(h0, hk, c32, dt[i], mode=quad, precision="fp32")
                    assert rel_err(u, ref) < 1e-5, (seed, mode, i, dims[i])


def test_graph_replays_and_loop_agree():
    from parament_amd.ensemble import Ensemble
    dt = [0.02] * len(DIMS)
    hs, cs = make_ensemble(8, DIMS, AMPS, PTS, dt)
    with Ensemble(DIMS, AMPS, PTS, dt) as e:
        e.set_hamiltonians(hs)
        e.set_execution("loop")
        loop = e.equiprop(cs)
        e.set_execution("graph")
        first = e.equiprop(cs)
        cfg = e.tuning
        assert cfg[16][0] == 1 and cfg[32][0] in (1, 2) and cfg[64][0] in (1, 2)
        # large instances run on four branches, or one per instance when fewer
        n_large = int(np.sum(np.asarray(DIMS) > 64))
        assert all(cfg[k][1] in (2, 3) for k in (16, 32, 64)) and cfg["branches"] == max(1, min(4, n_large))
        for _ in range(3):
            np.sum(np.random.default_rng(0).normal(size=1000))   # host work between replays
            again = e.equiprop(cs)
            for a, b in zip(first, again):
                assert np.array_equal(a, b)
        for a, b in zip(first, loop):
            assert rel_err(a, b) < 1e-5
        # new controls through the same graph
        cs2 = [c[::-1, ::-1].copy() for c in cs]
        e.set_execution("loop")
        loop2 = e.equiprop(cs2)
        e.set_execution("graph")
        for a, b in zip(e.equiprop(cs2), loop2):
            assert rel_err(a, b) < 1e-5


def test_ensemble_errors():
    from parament_amd.ensemble import Ensemble
    with pytest.raises(ValueError):
        Ensemble([4, 0], [1, 1], [10, 10], 0.1)
    with Ensemble([4], [1], [10], 0.1) as e:
        with pytest.raises(RuntimeError, match="No hamiltonian"):
            e.equiprop([np.zeros((1, 10))])
        with pytest.raises(ValueError):
            e.set_hamiltonians([[np.eye(4)]])
    with Ensemble([4], [1], [10], 100.0) as e:   # tau |H| beyond the iteration table
        rng = np.random.default_rng(1)
        with pytest.raises(RuntimeError, match="Timestep too large"):
            e.set_hamiltonians([[gue(rng, 4) * 50, gue(rng, 4)]])


GDIMS = [3, 16, 32, 33, 48, 70, 5]
GAMPS = [2, 1, 3, 2, 1, 2, 4]
GPTS = [21, 40, 17, 9, 23, 11, 64]


def _targets(seed, dims):
    from parament_amd.workloads import haar_unitary
    rng = np.random.default_rng(seed)
    return [haar_unitary(rng, d) for d in dims]


@pytest.mark.parametrize("quad", ["none", "midpoint", "simpson"])
def test_ensemble_gradient_matches_block_expm_oracle(quad):
    """Every instance's F and dF/dc (fused kernels up to dim 32, batched GEMMs above) against
    the block-matrix-expm gradient of its own problem, to complex64 accuracy."""
    from oracle import grape_oracle as go
    from parament_amd.ensemble import Ensemble
    pts = [p + 2 for p in GPTS] if quad == "simpson" else GPTS
    dt = [0.03 + 0.01 * i for i in range(len(GDIMS))]
    hs, cs = make_ensemble(11, GDIMS, GAMPS, pts, dt)
    ts = _targets(12, GDIMS)
    with Ensemble(GDIMS, GAMPS, pts, dt, quad) as e:
        e.set_hamiltonians(hs)
        fid, grads = e.gate_fidelity_gradient(cs, ts)
    for i, (h, c, t) in enumerate(zip(hs, cs, ts)):
        f_ref, g_ref = go.fidelity_gradient(h[0], h[1:], c, dt[i], quad, t)
        assert abs(fid[i] - f_ref) < 1e-5 * max(f_ref, 1e-3), (i, GDIMS[i])
        assert grads[i].shape == c.shape and grads[i].dtype == np.float32
        err = np.abs(grads[i] - g_ref).max() / np.abs(g_ref).max()
        assert err < 1e-4, (i, GDIMS[i], err)


def test_ensemble_gradient_graph_loop_and_context_agree():
    """The captured graph equals the per-instance loop bit for bit, replays repeat it, new
    controls go through the same graph, and each instance matches a complex64 context."""
    import parament_amd as parament
    from parament_amd.ensemble import Ensemble
    dt = [0.05] * len(GDIMS)
    hs, cs = make_ensemble(13, GDIMS, GAMPS, GPTS, dt)
    ts = _targets(14, GDIMS)
    with Ensemble(GDIMS, GAMPS, GPTS, dt) as e:
        e.set_hamiltonians(hs)
        e.set_execution("loop")
        f_loop, g_loop = e.gate_fidelity_gradient(cs, ts)
        e.set_execution("graph")
        f_graph, g_graph = e.gate_fidelity_gradient(cs, ts)
        assert np.array_equal(f_loop, f_graph)
        for a, b in zip(g_loop, g_graph):
            assert np.array_equal(a, b)
        for _ in range(2):
            np.sum(np.random.default_rng(0).normal(size=1000))   # host work between replays
            f2, g2 = e.gate_fidelity_gradient(cs, ts)
            assert np.array_equal(f2, f_graph)
            for a, b in zip(g2, g_graph):
                assert np.array_equal(a, b)
        cs2 = [-c for c in cs]
        f3, g3 = e.gate_fidelity_gradient(cs2, ts)
        e.set_execution("loop")
        f4, g4 = e.gate_fidelity_gradient(cs2, ts)
        assert np.array_equal(f3, f4)
        for a, b in zip(g3, g4):
            assert np.array_equal(a, b)
        assert not np.array_equal(g3[0], g_graph[0])
    for i, (h, c, t) in enumerate(zip(hs, cs, ts)):
        with parament.Parament("fp32") as p:
            p.set_hamiltonian(h[0], *h[1:])
            f, g = p.gate_fidelity_gradient(dt[i], *c, target=t)
        assert abs(f - f_graph[i]) <= 1e-6 * max(f, 1e-3)
        assert np.abs(g - g_graph[i]).max() <= 1e-6 * np.abs(g).max(), i


def test_ensemble_gradient_device_pointers():
    import torch
    from parament_amd.ensemble import Ensemble
    dt = [0.04] * len(GDIMS)
    hs, cs = make_ensemble(15, GDIMS, GAMPS, GPTS, dt)
    ts = _targets(16, GDIMS)
    dev = torch.device("cuda:0")
    carr = torch.tensor(np.concatenate([np.ravel(c) for c in cs]), dtype=torch.complex64, device=dev)
    tgt = torch.tensor(np.concatenate([np.ravel(t) for t in ts]), dtype=torch.complex64, device=dev)
    fid = torch.zeros(len(GDIMS), dtype=torch.float32, device=dev)
    grad = torch.zeros(carr.numel(), dtype=torch.float32, device=dev)
    with Ensemble(GDIMS, GAMPS, GPTS, dt) as e:
        e.set_hamiltonians(hs)
        f_host, g_host = e.gate_fidelity_gradient(cs, ts)
        e.gate_fidelity_gradient_device(carr.data_ptr(), tgt.data_ptr(), fid.data_ptr(), grad.data_ptr())
        e.synchronize()
    assert np.array_equal(fid.cpu().numpy(), f_host)
    assert np.array_equal(grad.cpu().numpy(), np.concatenate([np.ravel(g) for g in g_host]))


def test_ensemble_gradient_errors():
    from parament_amd.ensemble import Ensemble
    with Ensemble([4], [1], [10], 0.1) as e:
        with pytest.raises(RuntimeError, match="No hamiltonian"):
            e.gate_fidelity_gradient([np.zeros((1, 10))], [np.eye(4)])
        e.set_hamiltonians([[np.eye(4), np.eye(4)]])
        with pytest.raises(ValueError):
            e.gate_fidelity_gradient([np.zeros((1, 10)) + 1j], [np.eye(4)])
        with pytest.raises(ValueError):
            e.gate_fidelity_gradient([np.zeros((1, 10))], [np.eye(3)])


def test_full_size_ensemble512_forward_and_gradient():
    """The ensemble workload at full size (512 systems, dims 4-128, 1-4 controls, 256-8192 steps;
    one tuned, captured graph): every propagator unitary to complex64 accuracy, and the smallest,
    the largest and two mid-sized systems equal a complex64 context's propagator and gradient."""
    import parament_amd as parament
    from parament_amd import workloads
    from parament_amd.ensemble import Ensemble
    dims, amps, pts, dt, hs, cs = workloads.ensemble512()
    rng = np.random.default_rng(21)
    ts = [workloads.haar_unitary(rng, int(d)) for d in dims]
    with Ensemble(dims, amps, pts, dt) as e:
        e.set_hamiltonians(hs)
        us = e.equiprop(cs)
        fid, grads = e.gate_fidelity_gradient(cs, ts)
    for i, u in enumerate(us):
        d, n = int(dims[i]), int(pts[i])
        # complex64 rounding of n slices adding up as a random walk, with a wide margin
        assert np.linalg.norm(u.conj().T @ u - np.eye(d)) < 1e-6 * d * np.sqrt(n), (i, d, n)
    order = np.argsort(dims, kind="stable")
    for i in (order[0], order[-1], order[len(order) // 3], order[2 * len(order) // 3]):
        with parament.Parament("fp32") as p:
            p.set_hamiltonian(hs[i][0], *hs[i][1:])
            u = p.equiprop(dt, *cs[i])
            f, g = p.gate_fidelity_gradient(dt, *cs[i], target=ts[i])
        assert rel_err(us[i], u) < 1e-5, (i, int(dims[i]))
        assert abs(fid[i] - f) <= 1e-5 * max(f, 1e-3), (i, int(dims[i]))
        assert np.abs(grads[i] - g).max() <= 1e-4 * np.abs(g).max(), (i, int(dims[i]))


@pytest.mark.parametrize("quad,pts", [("none", 1), ("none", 2), ("midpoint", 2), ("simpson", 3), ("simpson", 4)])
def test_ensemble_gradient_few_samples(quad, pts):
    """One or two slices per system (no tree above the leaves, or one product), on the fused
    (dims 3, 20) and batched-GEMM (dim 40) paths."""
    from oracle import grape_oracle as go
    from parament_amd.ensemble import Ensemble
    dims, amps = [3, 20, 40], [2, 1, 2]
    ptss = [pts] * len(dims)
    dt = [0.05] * len(dims)
    hs, cs = make_ensemble(31 + pts, dims, amps, ptss, dt)
    ts = _targets(32 + pts, dims)
    with Ensemble(dims, amps, ptss, dt, quad) as e:
        e.set_hamiltonians(hs)
        fid, grads = e.gate_fidelity_gradient(cs, ts)
    for i, (h, c, t) in enumerate(zip(hs, cs, ts)):
        f_ref, g_ref = go.fidelity_gradient(h[0], h[1:], c, dt[i], quad, t)
        assert abs(fid[i] - f_ref) < 1e-5 * max(f_ref, 1e-3), (i, dims[i])
        assert np.abs(grads[i] - g_ref).max() <= 1e-4 * max(np.abs(g_ref).max(), 1e-3), (i, dims[i])
