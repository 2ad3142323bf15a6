"""The ``Parament`` context class: drop-in for pyparament's
(reference: src/python/pyparament/parament/parament.py:48-315).

Same constructor, methods, argument conventions, layouts, exception types and
messages.  Differences, all additive or fixes:

* ``dt`` is passed as a C double for both precisions (the reference's fp32 branch uses
  ``np.float``, which numpy >= 1.24 no longer has);
* error messages are read with ``Parament_errorMessage`` of the returned code (the reference
  calls a symbol its library does not export);
* ``set_iteration_cycles`` / ``automatic_iteration_cycles`` bind the C API's manual
  Chebyshev order (the reference binds neither), ``iteration_cycles`` reports the order
  used by the last propagation, and ``set_precision_mode`` selects the arithmetic;
* a context that is garbage-collected while the interpreter shuts down is left to the
  process exit instead of calling into the library.
"""
import ctypes
import sys

import numpy as np

from .constants import *  # noqa: F401,F403
from .constants import (PARAMENT_FAIL, PARAMENT_PRECISION_MIXED, PARAMENT_PRECISION_MIXED_BF16X2,
                        PARAMENT_PRECISION_MIXED_FP32, PARAMENT_PRECISION_NATIVE,
                        PARAMENT_QUADRATURE_MIDPOINT, PARAMENT_QUADRATURE_NONE, PARAMENT_QUADRATURE_SIMPSON,
                        PARAMENT_STATUS_CUBLAS_FAILED, PARAMENT_STATUS_CUBLAS_INIT_FAILED,
                        PARAMENT_STATUS_DEVICE_ALLOC_FAILED, PARAMENT_STATUS_HOST_ALLOC_FAILED,
                        PARAMENT_STATUS_INVALID_QUADRATURE_SELECTION, PARAMENT_STATUS_INVALID_VALUE,
                        PARAMENT_STATUS_NO_HAMILTONIAN, PARAMENT_STATUS_SELECT_SMALLER_DT, PARAMENT_STATUS_SUCCESS,
                        PARAMENT_STATUS_WORKSPACE_TOO_SMALL)
from . import paramentlib

try:
    import qutip
except ImportError:
    from . import qutip_mock as qutip

_QUADRATURE = {
    'none': PARAMENT_QUADRATURE_NONE,
    'midpoint': PARAMENT_QUADRATURE_MIDPOINT,
    'simpson': PARAMENT_QUADRATURE_SIMPSON,
}

_PRECISION = {
    'native': PARAMENT_PRECISION_NATIVE,
    'mixed': PARAMENT_PRECISION_MIXED,
    'mixed_fp32': PARAMENT_PRECISION_MIXED_FP32,
    'mixed_bf16x2': PARAMENT_PRECISION_MIXED_BF16X2,
}

# Exception class raised for each C status code (the reference's mapping: allocation failures
# are MemoryError, rejected arguments ValueError, everything else RuntimeError).
_STATUS_EXCEPTION = {
    PARAMENT_STATUS_HOST_ALLOC_FAILED: MemoryError,
    PARAMENT_STATUS_DEVICE_ALLOC_FAILED: MemoryError,
    PARAMENT_STATUS_WORKSPACE_TOO_SMALL: MemoryError,
    PARAMENT_STATUS_INVALID_VALUE: ValueError,
    PARAMENT_STATUS_INVALID_QUADRATURE_SELECTION: ValueError,
}
for _code in (PARAMENT_STATUS_CUBLAS_INIT_FAILED, PARAMENT_STATUS_CUBLAS_FAILED, PARAMENT_STATUS_SELECT_SMALLER_DT,
              PARAMENT_STATUS_NO_HAMILTONIAN, PARAMENT_FAIL):
    _STATUS_EXCEPTION[_code] = RuntimeError
del _code


def status_exception(lib, code):
    """The exception a non-zero status code stands for, carrying the library's message as
    ``"Error code <code>: <message>"``.  A code the C API does not define is a broken
    invariant between this module and the library, reported as AssertionError."""
    kind = _STATUS_EXCEPTION.get(code)
    if kind is None:
        return AssertionError(f'Unknown error code {code}')
    return kind(f"Error code {code}: {lib.Parament_errorMessage(code).decode()}")


def _dense(h):
    """An ndarray view of a Hamiltonian given as ndarray, nested list or qutip object."""
    if isinstance(h, qutip.Qobj):
        data = h.data
        return np.asarray(data.todense() if hasattr(data, 'todense') else h.full())
    return h


def _control_block(carr, n_hamiltonians):
    """The control waveforms of an equiprop call as one (amps, pts) array, after the checks the
    reference makes: no more arrays than control Hamiltonians, all of them one-dimensional and
    of one length."""
    if len(carr) > n_hamiltonians:
        raise ValueError(f'Got {len(carr)} amplitude arrays, but there are only {n_hamiltonians} Hamiltonians.')
    if not carr:
        raise ValueError('provide at least 1 control amplitude')
    shapes = {np.shape(waveform) for waveform in carr}
    if len(shapes) != 1 or len(next(iter(shapes))) != 1:
        raise ValueError("All amplitude arrays must have the same length.")
    return np.stack([np.asarray(waveform) for waveform in carr])


class Parament:
    """A Parament context: settings and device state shared by successive calls.

    Load the Hamiltonians with :meth:`set_hamiltonian`, then compute propagators with
    :meth:`equiprop`.

    Parameters
    ----------
    precision : {'fp32', 'fp64'}
        Native precision of the context (complex64 or complex128 arrays).

    Examples
    --------
    >>> gpu_runner = Parament()
    >>> H0 = np.array([[1, 0], [0, -1]])
    >>> H1 = np.array([[0, 1], [1, 0]])
    >>> gpu_runner.set_hamiltonian(H0, H1)
    >>> gpu_runner.equiprop(1.0, np.zeros(1))
    array([[0.5403023-0.84147096j, 0.       +0.j        ],
           [0.       +0.j        , 0.5403023+0.84147096j]], dtype=complex64)

    Use the context as a context manager (``with Parament() as p: ...``) or call
    :meth:`destroy` to release its resources deterministically.
    """

    def __init__(self, precision='fp32'):
        if precision not in ('fp32', 'fp64'):
            raise ValueError("precision must be either 'fp32' or 'fp64'")
        self._use_doubles = precision == 'fp64'
        self._sfx = '_fp64' if self._use_doubles else ''
        self._ctype = np.complex128 if self._use_doubles else np.complex64
        self._lib = paramentlib.lib
        # bound now, so that __del__ needs no module globals during interpreter shutdown
        self._c_destroy = getattr(self._lib, 'Parament_destroy' + self._sfx)
        self._shutting_down = sys.is_finalizing
        self._handle = None
        handle = ctypes.c_void_p()
        self._ok(self._call('Parament_create', ctypes.byref(handle)))
        self._handle = handle
        self.dim = -1
        self.amps = -1
        self._qutip_H0 = None
        self._use_qutip = False

    # ---- plumbing ----

    def _call(self, name, *args):
        return getattr(self._lib, name + self._sfx)(*args)

    def _ok(self, code):
        """Raise the reference's exception for a failed C call."""
        if code != PARAMENT_STATUS_SUCCESS:
            raise status_exception(self._lib, code)

    def _live(self):
        """The C handle; a destroyed context raises instead."""
        if self._handle is None:
            raise RuntimeError("Attempting to use a context that has been destroyed")
        return self._handle

    def _get_error_message(self, code=None):
        """Message of `code`, or of the context's last status when `code` is None."""
        if code is None:
            code = self._call('Parament_peekAtLastError', self._live())
        return self._lib.Parament_errorMessage(code).decode()

    # ---- reference API ----

    def destroy(self):
        """Destroy the context; also happens when the object is garbage-collected."""
        handle = self._live()
        self._handle = None
        self._ok(self._c_destroy(handle))

    def __del__(self):
        handle = getattr(self, '_handle', None)
        if handle is None:
            return
        self._handle = None
        try:
            if not self._shutting_down():
                self._c_destroy(handle)
        except Exception:  # noqa: BLE001 - nothing can be reported from a finaliser
            pass

    def set_hamiltonian(self, H0, *H1, use_magnus=False, quadrature_mode='none'):
        """Load the drift Hamiltonian ``H0`` and the control Hamiltonians ``*H1``.

        ``H0`` may be an ndarray or a qutip object (then :meth:`equiprop` returns qutip
        objects too).  ``H1`` is variadic; a single 3D array whose first axis enumerates
        the Hamiltonians also works.  ``use_magnus`` requires ``quadrature_mode='simpson'``.
        """
        handle = self._live()
        self._use_qutip = isinstance(H0, qutip.Qobj)
        self._qutip_H0 = H0 if self._use_qutip else None
        drift = np.atleast_2d(_dense(H0))
        if quadrature_mode not in _QUADRATURE:
            raise ValueError('unknown quadrature mode selected')
        if not H1:
            raise ValueError('provide at least 1 control amplitude')
        controls = np.atleast_2d([_dense(h) for h in H1])
        # a 2-D stack means one control Hamiltonian; 3-D stacks enumerate them along axis 0
        n_controls = controls.shape[0] if controls.ndim > 2 else 1
        dim = drift.shape[0]
        self.dim, self.amps = dim, n_controls
        # Row-major hand-over: the C side then works on transposes, which keeps the time order of
        # the product without reversing the controls (parament.py:197-198).
        flat = lambda a: np.ascontiguousarray(a, dtype=self._ctype).reshape(-1)  # noqa: E731
        self._ok(self._call('Parament_setHamiltonian', handle, flat(drift), flat(controls), dim, n_controls,
                            bool(use_magnus), _QUADRATURE[quadrature_mode]))

    def equiprop(self, dt, *carr):
        """Propagator of the loaded Hamiltonians under the control waveforms ``*carr``.

        One 1D array of samples per control Hamiltonian (fewer arrays than Hamiltonians
        leave the rest at zero amplitude); a 2D array unpacks with ``*``.
        Returns a ``(dim, dim)`` ndarray (or a qutip object).
        """
        handle = self._live()
        if self.amps < 0:
            raise RuntimeError("No hamiltonian set")
        block = _control_block(carr, self.amps)
        n_fields, n_pts = block.shape
        result = np.empty((self.dim, self.dim), dtype=self._ctype)
        samples = np.ascontiguousarray(block, dtype=self._ctype).reshape(-1)
        self._ok(self._call('Parament_equiprop', handle, samples, float(dt), n_pts, n_fields, result.reshape(-1)))
        if self._use_qutip:
            return qutip.Qobj(result, dims=self._qutip_H0.dims)
        return result

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        self.destroy()

    # ---- additions (the C API has them; the reference's Python does not bind them) ----

    def set_iteration_cycles(self, cycles):
        """Force the Chebyshev order M (odd, >= 3); see Parament_setIterationCyclesManually."""
        self._ok(self._call('Parament_setIterationCyclesManually', self._live(), int(cycles)))

    def automatic_iteration_cycles(self):
        """Restore the automatic choice of M from the iteration tables."""
        self._ok(self._call('Parament_automaticIterationCycles', self._live()))

    @property
    def iteration_cycles(self):
        """Chebyshev order M used by the last successful :meth:`equiprop` (-1 before any)."""
        if self._handle is None:
            return -1
        return self._call('Parament_getIterationCycles', self._handle)

    def set_precision_mode(self, mode):
        """Arithmetic of the context's equiprop.

        'native': the context's own precision throughout (the reference's arithmetic; default of
        'fp64' contexts).  'mixed': FP64 for the norm, the Chebyshev coefficients, the slice
        Hamiltonians, the first-order terms and the upper product-tree levels, FP32-accurate
        split products for the rest: scaled fp16 two-way splits where the slice polynomial is
        evaluated in powers of tau X (tau Hnorm <= 1), bf16 three-way splits for the recurrence
        (default of 'fp32' contexts, which keep the wavefront-resident FP32 kernels for dims
        <= 16).  'mixed_fp32' and 'mixed_bf16x2': the same with FP32 or bf16 hi+lo products.
        gate_fidelity_gradient of an 'fp32' context above dim 32 runs its products on the fp16
        splits in 'mixed' (power form) and on FP32 products in the other modes."""
        code = _PRECISION.get(mode)
        if code is None:
            raise ValueError("mode must be 'native', 'mixed', 'mixed_fp32' or 'mixed_bf16x2'")
        self._ok(self._call('Parament_setPrecisionMode', self._live(), code))

    @property
    def precision_mode(self):
        """The current arithmetic as a :meth:`set_precision_mode` name."""
        code = self._call('Parament_getPrecisionMode', self._live())
        return {v: k for k, v in _PRECISION.items()}[code]

    def set_kernel_path(self, path):
        """'auto' (fused on-chip kernels for dim <= 64) or 'stepwise' (batched step kernels)."""
        code = {'auto': 0, 'stepwise': 1}.get(path)
        if code is None:
            raise ValueError("path must be 'auto' or 'stepwise'")
        self._ok(self._call('Parament_setKernelPath', self._live(), code))

    def equiprop_device(self, dt, carr_ptr, pts, amps, out_ptr):
        """Asynchronous equiprop on device pointers (controls and result stay in HBM)."""
        self._ok(self._call('Parament_equipropDevice', self._live(), ctypes.c_void_p(carr_ptr), float(dt), int(pts),
                            int(amps), ctypes.c_void_p(out_ptr)))

    def gate_fidelity_gradient(self, dt, *carr, target):
        """F = |Tr(target^H U)|^2 / dim^2 and dF/dc for every control sample (adjoint method).

        Real controls, any dim, every quadrature mode including Magnus.  Returns ``(F, grad)`` with ``grad`` shaped like
        the stacked controls, ``(amps, pts)``, in the context's precision.
        """
        handle = self._live()
        if self.amps < 0:
            raise RuntimeError("No hamiltonian set")
        block = _control_block(carr, self.amps)
        if np.iscomplexobj(block) and np.any(block.imag != 0):
            raise ValueError("gradients need real control amplitudes")
        n_fields, n_pts = block.shape
        rtype = np.float64 if self._use_doubles else np.float32
        samples = np.ascontiguousarray(block.real, dtype=self._ctype).reshape(-1)
        tgt = np.ascontiguousarray(target, dtype=self._ctype).reshape(-1)
        if tgt.size != self.dim ** 2:
            raise ValueError("target must be a dim x dim matrix")
        fid = ctypes.c_double() if self._use_doubles else ctypes.c_float()
        grad = np.zeros(n_fields * n_pts, dtype=rtype)
        self._ok(self._call('Parament_gateFidelityGradient', handle, samples, float(dt), n_pts, n_fields, tgt,
                            ctypes.byref(fid), grad))
        return fid.value, grad.reshape(n_fields, n_pts)

    def gate_fidelity_gradient_device(self, dt, carr_ptr, pts, amps, target_ptr, fid_ptr, grad_ptr):
        """Device-pointer variant (asynchronous on the context's stream); reads the real parts of
        the complex control array at ``carr_ptr``."""
        self._ok(self._call('Parament_gateFidelityGradientDevice', self._live(), ctypes.c_void_p(carr_ptr), float(dt),
                            int(pts), int(amps), ctypes.c_void_p(target_ptr),
                            ctypes.c_void_p(fid_ptr) if fid_ptr else None, ctypes.c_void_p(grad_ptr)))

    def set_gradient_checkpointing(self, segment=1, chunk=0):
        """Gradient memory policy: ``segment=0`` keeps the whole product tree (default of a new
        context); ``segment=1`` checkpoints at ~sqrt(N)-slice segments; larger values set the
        segment length.  ``chunk`` is the number of slices recomputed per pass (0: automatic)."""
        self._ok(self._call('Parament_setGradientCheckpointing', self._live(), int(segment), int(chunk)))

    @property
    def gradient_workspace(self):
        """Device bytes the last gradient call worked in."""
        return int(self._call('Parament_getGradientWorkspace', self._live()))

    def set_stream(self, stream_ptr):
        """Run on the given HIP stream (an int handle, e.g. torch.cuda.current_stream().cuda_stream).

        0 / None is the device's null stream (PyTorch's default stream); passing back a value
        read from ``stream`` before restores the context's own stream."""
        self._ok(self._call('Parament_setStream', self._live(), ctypes.c_void_p(stream_ptr or None)))

    @property
    def stream(self):
        return self._call('Parament_getStream', self._handle) if self._handle is not None else None

    def set_timing(self, enable=True):
        self._ok(self._call('Parament_setTiming', self._live(), int(bool(enable))))

    def kernel_times(self):
        """(slice-kernel ms, product-tree ms) of the last equiprop, with timing enabled."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._ok(self._call('Parament_getKernelTimes', self._live(), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def synchronize(self):
        self._ok(self._call('Parament_synchronize', self._live()))
