"""Allocation-free equiprop (Parament_equipropDeviceNoAlloc_fp64) through the C ABI against the
existing device entry point (Parament_equipropDevice_fp64): the same kernels on the same
inputs, so the propagators agree to the complex128 bar (1e-12, in practice bit for bit), on two
calls in a row of which the second is smaller, for each precision mode and kernel family; and
a problem beyond the fixed workspace is refused with WORKSPACE_TOO_SMALL."""
import ctypes

import numpy as np
import pytest

from oracle import equiprop_oracle as eo
from tests.helpers import gue, rel_err

pytestmark = pytest.mark.gpu

WORKSPACE_TOO_SMALL = 100


def _device_calls(p, c_list, dt):
    """For each control array: (result of equipropDevice, result of the allocation-free call)."""
    import torch
    lib = p._lib
    d = p.dim
    out = []
    for c in c_list:
        amps, pts = c.shape
        carr = torch.tensor(np.ascontiguousarray(c), dtype=torch.complex128, device="cuda:0")
        u_ref = torch.full((d * d,), float("nan"), dtype=torch.complex128, device="cuda:0")
        u_new = torch.full((d * d,), float("nan"), dtype=torch.complex128, device="cuda:0")
        torch.cuda.synchronize()
        assert lib.Parament_equipropDevice_fp64(p._handle, ctypes.c_void_p(carr.data_ptr()), dt, pts, amps,
                                                ctypes.c_void_p(u_ref.data_ptr())) == 0
        assert lib.Parament_synchronize_fp64(p._handle) == 0
        assert lib.Parament_equipropDeviceNoAlloc_fp64(p._handle, ctypes.c_void_p(carr.data_ptr()), dt, pts, amps,
                                                       ctypes.c_void_p(u_new.data_ptr())) == 0
        assert lib.Parament_synchronize_fp64(p._handle) == 0
        out.append((u_ref.cpu().numpy().reshape(d, d), u_new.cpu().numpy().reshape(d, d)))
    return out


@pytest.mark.parametrize("d,mode,path,quad,pts", [
    (8, "native", "auto", "none", 2049),          # wavefront-resident small-dim kernel + tree
    (48, "native", "auto", "simpson", 2049),      # fused FP64 slice kernel + tree
    (64, "mixed", "auto", "simpson", 8193),       # fused mixed kernel (two folded levels) + FP64 tree
    (20, "mixed_bf16x2", "auto", "none", 2049),   # batched split-product GEMMs
    (72, "native", "auto", "midpoint", 257),      # batched FP64 GEMMs (3 x 256 slices of 83 KB)
    (24, "mixed", "stepwise", "simpson", 2049),   # batched GEMMs, mixed, forced
])
def test_no_alloc_matches_device_entry_point(d, mode, path, quad, pts):
    import parament_amd as parament
    rng = np.random.default_rng(9000 + d)
    h = [gue(rng, d) for _ in range(3)]
    c = rng.uniform(-1, 1, (2, pts))
    small = c[:, :pts // 7]
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(h[0], *h[1:], quadrature_mode=quad)
        p.set_precision_mode(mode)
        p.set_kernel_path(path)
        # two calls in a row, the second smaller (the workspace keeps the first call's contents)
        res = _device_calls(p, [c, small], 0.01)
        for (u_ref, u_new), cc in zip(res, (c, small)):
            assert np.all(np.isfinite(u_new))
            assert rel_err(u_new, u_ref) <= 1e-12
            if mode == "native":
                assert rel_err(u_new, eo.equiprop(h[0], h[1:], cc, 0.01, mode=quad)) < 1e-12


def test_no_alloc_headline_size_fits_and_large_problems_are_refused():
    """Dim 64 at 65,536 steps in the mixed mode fits the 256 MiB workspace; 32,767 FP64 slices at
    dim 64 (2 GiB of slices) do not, and the call fails cleanly without touching the output."""
    import torch
    import parament_amd as parament
    from parament_amd import workloads
    h0, hk, c, dt, quad = workloads.heisenberg6()
    with parament.Parament("fp64") as p:
        p.set_hamiltonian(h0, *hk, quadrature_mode=quad)
        p.set_precision_mode("mixed")
        (u_ref, u_new), = _device_calls(p, [c], dt)
        assert rel_err(u_new, u_ref) <= 1e-12
        p.set_precision_mode("native")
        carr = torch.tensor(np.ascontiguousarray(c), dtype=torch.complex128, device="cuda:0")
        out = torch.zeros(64 * 64, dtype=torch.complex128, device="cuda:0")
        code = p._lib.Parament_equipropDeviceNoAlloc_fp64(p._handle, ctypes.c_void_p(carr.data_ptr()), dt,
                                                          c.shape[1], c.shape[0], ctypes.c_void_p(out.data_ptr()))
        assert code == WORKSPACE_TOO_SMALL
        assert p._lib.Parament_peekAtLastError_fp64(p._handle) == WORKSPACE_TOO_SMALL
        assert "fixed workspace" in p._lib.Parament_errorMessage(code).decode()
        assert p._lib.Parament_synchronize_fp64(p._handle) == 0
        assert torch.count_nonzero(out).item() == 0
        # the context is still usable afterwards, through either entry point
        (u_ref, u_new), = _device_calls(p, [c[:, :1025]], dt)
        assert rel_err(u_new, u_ref) <= 1e-12
