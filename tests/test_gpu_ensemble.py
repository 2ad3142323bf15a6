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
        assert all(cfg[k][1] in (2, 3) for k in (16, 32, 64)) and cfg["branches"] in (1, 2, 4)
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
