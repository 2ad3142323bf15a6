"""Ensembles of independent complex64 equiprop problems (Parament_ensemble* in libparament).

No counterpart in the reference package: many systems of different dimension, control count
and length solved together.  Each instance gets what ``Parament('fp32').equiprop`` would return
for it (same iteration-count table, quadrature and product order; FP32 products up to dim 64,
FP32-accurate fp16 split products above, as the default mode's step-wise path); the ensemble
runs as one captured HIP graph with multi-instance kernels for the small dims, batched GEMMs for
dims above 64, concurrent branches and tuned tree block sizes and grids.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import paramentlib
from .constants import (PARAMENT_QUADRATURE_MIDPOINT, PARAMENT_QUADRATURE_NONE, PARAMENT_QUADRATURE_SIMPSON,
                        PARAMENT_STATUS_SUCCESS)
from .parament import _dense, status_exception

__all__ = ["Ensemble"]

_QUAD = {"none": PARAMENT_QUADRATURE_NONE, "midpoint": PARAMENT_QUADRATURE_MIDPOINT,
         "simpson": PARAMENT_QUADRATURE_SIMPSON}


class Ensemble:
    """Fixed-shape ensemble: ``dims[i]`` x ``dims[i]`` Hamiltonians with ``amps[i]`` controls
    of ``pts[i]`` samples and time step ``dt[i]``; one quadrature mode for all."""

    def __init__(self, dims, amps, pts, dt, quadrature_mode="none"):
        self._lib = paramentlib.lib
        self.dims = np.ascontiguousarray(dims, dtype=np.uint32)
        self.amps = np.ascontiguousarray(amps, dtype=np.uint32)
        self.pts = np.ascontiguousarray(pts, dtype=np.uint32)
        self.dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), self.dims.shape))
        if not (self.dims.shape == self.amps.shape == self.pts.shape) or self.dims.ndim != 1:
            raise ValueError("dims, amps and pts must be 1-d sequences of the same length")
        if quadrature_mode not in _QUAD:
            raise ValueError(f"Unknown quadrature mode {quadrature_mode!r}")
        self._handle = ctypes.c_void_p()
        self._check(self._lib.Parament_ensembleCreate(ctypes.byref(self._handle), len(self.dims), self.dims, self.amps,
                                                      self.pts, self.dt, _QUAD[quadrature_mode]))

    def _check(self, code):
        if code != PARAMENT_STATUS_SUCCESS:
            raise status_exception(self._lib, code)

    def set_hamiltonians(self, hamiltonians):
        """``hamiltonians[i]`` = (H0, H1, ..., H_amps[i]) for instance i."""
        if len(hamiltonians) != len(self.dims):
            raise ValueError("one Hamiltonian set per instance")
        parts = []
        for i, hs in enumerate(hamiltonians):
            hs = [np.asarray(_dense(h), dtype=np.complex64) for h in hs]
            d = int(self.dims[i])
            if len(hs) != 1 + int(self.amps[i]) or any(h.shape != (d, d) for h in hs):
                raise ValueError(f"instance {i}: expected {1 + int(self.amps[i])} matrices of shape ({d}, {d})")
            parts += [np.ravel(h, order="C") for h in hs]
        self._check(self._lib.Parament_ensembleSetHamiltonians(self._handle, np.ascontiguousarray(np.concatenate(parts))))

    def _pack_controls(self, controls):
        if len(controls) != len(self.dims):
            raise ValueError("one control array per instance")
        parts = []
        for i, c in enumerate(controls):
            c = np.asarray(c)
            if c.shape != (int(self.amps[i]), int(self.pts[i])):
                raise ValueError(f"instance {i}: controls must have shape ({int(self.amps[i])}, {int(self.pts[i])})")
            parts.append(np.ravel(c.astype(np.complex64), order="C"))
        return np.ascontiguousarray(np.concatenate(parts))

    def equiprop(self, controls):
        """Propagators of all instances (list of complex64 arrays)."""
        carr = self._pack_controls(controls)
        out = np.empty(int(np.sum(self.dims.astype(np.int64) ** 2)), dtype=np.complex64)
        self._check(self._lib.Parament_ensembleEquiprop(self._handle, carr, out))
        res, off = [], 0
        for d in self.dims.astype(int):
            res.append(out[off:off + d * d].reshape(d, d))
            off += d * d
        return res

    def gate_fidelity_gradient(self, controls, targets):
        """Per instance F_i = |Tr(targets[i]^H U_i)|^2 / d_i^2 and dF_i/dc (adjoint method, real
        controls): returns (array of F_i, list of (amps[i], pts[i]) float32 gradients)."""
        carr = self._pack_controls(controls)
        if np.any(carr.imag != 0):
            raise ValueError("gradients need real control amplitudes")
        if len(targets) != len(self.dims):
            raise ValueError("one target per instance")
        parts = []
        for i, t in enumerate(targets):
            t = np.asarray(t, dtype=np.complex64)
            d = int(self.dims[i])
            if t.shape != (d, d):
                raise ValueError(f"instance {i}: target must have shape ({d}, {d})")
            parts.append(np.ravel(t, order="C"))
        tgt = np.ascontiguousarray(np.concatenate(parts))
        fid = np.zeros(len(self.dims), dtype=np.float32)
        grad = np.zeros(carr.size, dtype=np.float32)
        self._check(self._lib.Parament_ensembleGateFidelityGradient(self._handle, carr, tgt, fid, grad))
        res, off = [], 0
        for a, p in zip(self.amps.astype(int), self.pts.astype(int)):
            res.append(grad[off:off + a * p].reshape(a, p))
            off += a * p
        return fid, res

    def gate_fidelity_gradient_device(self, carr_ptr, targets_ptr, fid_ptr, grad_ptr):
        """Device-pointer variant (packed layouts; fid_ptr may be 0); asynchronous."""
        self._check(self._lib.Parament_ensembleGateFidelityGradientDevice(
            self._handle, ctypes.c_void_p(carr_ptr), ctypes.c_void_p(targets_ptr),
            ctypes.c_void_p(fid_ptr) if fid_ptr else None, ctypes.c_void_p(grad_ptr)))

    def equiprop_device(self, carr_ptr, out_ptr):
        """Device-pointer variant on the ensemble's own stream (packed layouts); asynchronous."""
        self._check(self._lib.Parament_ensembleEquipropDevice(self._handle, ctypes.c_void_p(carr_ptr),
                                                              ctypes.c_void_p(out_ptr)))

    def synchronize(self):
        self._check(self._lib.Parament_ensembleSynchronize(self._handle))

    def set_execution(self, mode):
        """'graph' (default: multi-instance kernels in one captured graph) or 'loop' (one
        instance after the other)."""
        self._check(self._lib.Parament_ensembleSetExecution(self._handle, {"graph": 0, "loop": 1}[mode]))

    def tune(self):
        self._check(self._lib.Parament_ensembleTune(self._handle))

    @property
    def tuning(self):
        """{class dim: (T, log2 tree block, grid multiple)} and the number of branches."""
        cfg = np.zeros(10, dtype=np.int32)
        self._check(self._lib.Parament_ensembleGetTuning(self._handle, cfg))
        return {16: tuple(cfg[0:3]), 32: tuple(cfg[3:6]), 64: tuple(cfg[6:9]), "branches": int(cfg[9])}

    def destroy(self):
        if self._handle:
            self._lib.Parament_ensembleDestroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()
