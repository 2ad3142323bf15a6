"""pyparament extras on the GPU (reference parament.py:29-35,162-168,274-275; debug_functions.py:22-38;
deviceInfo.c:30-59): device_info's report, the qutip round trip, the debug expm helper."""
import types

import numpy as np
import pytest
import scipy.linalg

import parament_amd as parament
from parament_amd import qutip_mock
from tests.helpers import gue

pytestmark = pytest.mark.gpu


def test_device_info_reports_the_gpu(capfd):
    parament.device_info()
    out = capfd.readouterr().out
    assert "Number of devices: " in out
    n = int(out.split("Number of devices: ")[1].split()[0])
    assert n >= 1
    assert "gfx950" in out   # the architecture the library is built for
    assert "Compute units:" in out


class _Dense:
    """Sparse-matrix stand-in: what qutip's Qobj.data offers (todense)."""

    def __init__(self, a):
        self._a = a

    def todense(self):
        return np.asmatrix(self._a)


class DuckQobj(qutip_mock.Qobj):
    """A qutip-like object: .data with todense(), .full(), .dims, and the Qobj(data, dims=...)
    constructor the class uses to wrap its result."""

    def __init__(self, arr, dims=None):
        self._arr = np.asarray(arr)
        n = self._arr.shape[0]
        self.dims = dims if dims is not None else [[n], [n]]

    @property
    def data(self):
        return _Dense(self._arr)

    def full(self):
        return self._arr


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_qobj_in_qobj_out(monkeypatch, precision):
    monkeypatch.setattr(parament.parament, "qutip", types.SimpleNamespace(Qobj=DuckQobj))
    rng = np.random.default_rng(3)
    d = 4
    h0, h1 = gue(rng, d), gue(rng, d)
    c = rng.uniform(-1, 1, 33)
    dims = [[2, 2], [2, 2]]   # a two-qubit system: the result keeps the input's tensor structure
    with parament.Parament(precision) as ctx:
        ctx.set_hamiltonian(DuckQobj(h0, dims), DuckQobj(h1, dims), quadrature_mode="simpson")
        uq = ctx.equiprop(0.05, c)
        ctx.set_hamiltonian(h0, h1, quadrature_mode="simpson")
        ua = ctx.equiprop(0.05, c)
    assert isinstance(uq, DuckQobj) and uq.dims == dims
    assert isinstance(ua, np.ndarray)   # plain arrays in, plain arrays out
    np.testing.assert_array_equal(uq.full(), ua)
    tol = 1e-6 if precision == "fp32" else 1e-13
    assert np.linalg.norm(uq.full().conj().T @ uq.full() - np.eye(d)) < tol


def test_debug_expm_small_norm_matrices():
    rng = np.random.default_rng(7)
    for d in (2, 5, 12):
        m = (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))) * (0.3 / d)
        e = parament.expm(m)
        assert e.dtype == np.complex64
        assert np.linalg.norm(e - scipy.linalg.expm(m)) < 1e-6
