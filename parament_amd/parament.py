"""The ``Parament`` context class: drop-in for pyparament's
(reference: src/python/pyparament/parament/parament.py:48-315).

Same constructor, methods, argument conventions, layouts, exception types and
messages.  Differences, all additive or fixes:

* ``dt`` is passed as a C double for both precisions (the reference's fp32 branch uses
  ``np.float``, which numpy >= 1.24 no longer has);
* error messages are read with ``Parament_peekAtLastError`` (the reference calls a
  symbol its library does not export);
* ``set_iteration_cycles`` / ``automatic_iteration_cycles`` bind the C API's manual
  Chebyshev order (the reference binds neither), ``iteration_cycles`` reports the order
  used by the last propagation, and ``set_precision_mode`` selects the mixed-precision
  path of a complex128 context.
"""
import ctypes
import logging
from functools import wraps

import numpy as np

from .constants import *  # noqa: F401,F403
from .constants import (PARAMENT_FAIL, PARAMENT_PRECISION_MIXED, PARAMENT_PRECISION_MIXED_BF16X2,
                        PARAMENT_PRECISION_MIXED_FP32, PARAMENT_PRECISION_NATIVE,
                        PARAMENT_QUADRATURE_MIDPOINT, PARAMENT_QUADRATURE_NONE, PARAMENT_QUADRATURE_SIMPSON,
                        PARAMENT_STATUS_CUBLAS_FAILED, PARAMENT_STATUS_CUBLAS_INIT_FAILED,
                        PARAMENT_STATUS_DEVICE_ALLOC_FAILED, PARAMENT_STATUS_HOST_ALLOC_FAILED,
                        PARAMENT_STATUS_INVALID_QUADRATURE_SELECTION, PARAMENT_STATUS_INVALID_VALUE,
                        PARAMENT_STATUS_NO_HAMILTONIAN, PARAMENT_STATUS_SELECT_SMALLER_DT, PARAMENT_STATUS_SUCCESS)
from . import paramentlib

logger = logging.getLogger('Parament')
logger.setLevel(logging.CRITICAL)

try:
    import qutip
except ImportError:
    from . import qutip_mock as qutip

_QUADRATURE = {
    'none': PARAMENT_QUADRATURE_NONE,
    'midpoint': PARAMENT_QUADRATURE_MIDPOINT,
    'simpson': PARAMENT_QUADRATURE_SIMPSON,
}

_EXCEPTIONS = {
    PARAMENT_STATUS_HOST_ALLOC_FAILED: MemoryError,
    PARAMENT_STATUS_DEVICE_ALLOC_FAILED: MemoryError,
    PARAMENT_STATUS_CUBLAS_INIT_FAILED: RuntimeError,
    PARAMENT_STATUS_INVALID_VALUE: ValueError,
    PARAMENT_STATUS_CUBLAS_FAILED: RuntimeError,
    PARAMENT_STATUS_SELECT_SMALLER_DT: RuntimeError,
    PARAMENT_STATUS_NO_HAMILTONIAN: RuntimeError,
    PARAMENT_STATUS_INVALID_QUADRATURE_SELECTION: ValueError,
    PARAMENT_FAIL: RuntimeError,
}


def _must_be_alive(f):
    @wraps(f)
    def wrapper(self, *args, **kwargs):
        if self._handle is None:
            raise RuntimeError("Attempting to use a context that has been destroyed")
        return f(self, *args, **kwargs)
    return wrapper


def _dense(h):
    if isinstance(h, qutip.Qobj):
        data = h.data
        return np.asarray(data.todense() if hasattr(data, 'todense') else h.full())
    return h


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
        self._handle = ctypes.c_void_p()
        self._check_error(self._call('Parament_create', ctypes.byref(self._handle)))
        self.dim = -1
        self.amps = -1
        self._qutip_H0 = None
        self._use_qutip = False

    def _call(self, name, *args):
        return getattr(self._lib, name + self._sfx)(*args)

    @_must_be_alive
    def destroy(self):
        """Destroy the context; also happens when the object is garbage-collected."""
        self._check_error(self._call('Parament_destroy', self._handle))
        self._handle = None

    def __del__(self):
        if getattr(self, '_handle', None) is not None:
            self.destroy()

    @_must_be_alive
    def set_hamiltonian(self, H0, *H1, use_magnus=False, quadrature_mode='none'):
        """Load the drift Hamiltonian ``H0`` and the control Hamiltonians ``*H1``.

        ``H0`` may be an ndarray or a qutip object (then :meth:`equiprop` returns qutip
        objects too).  ``H1`` is variadic; a single 3D array whose first axis enumerates
        the Hamiltonians also works.  ``use_magnus`` requires ``quadrature_mode='simpson'``.
        """
        if isinstance(H0, qutip.Qobj):
            self._use_qutip = True
            self._qutip_H0 = H0
        else:
            self._use_qutip = False
        H0 = np.atleast_2d(_dense(H0))

        try:
            mode_sel = _QUADRATURE[quadrature_mode]
        except (KeyError, TypeError):
            raise ValueError('unknown quadrature mode selected') from None

        dim = np.shape(H0)[0]
        if len(H1) == 0:
            raise ValueError('provide at least 1 control amplitude')
        H1 = np.atleast_2d([_dense(Hi) for Hi in H1])
        shape = np.shape(H1)
        amps = shape[0] if len(shape) > 2 else 1
        self.dim = dim
        self.amps = amps
        # Row-major (C-order) hand-over, as the reference does (parament.py:197-212).
        self._check_error(self._call(
            'Parament_setHamiltonian', self._handle,
            np.ascontiguousarray(np.ravel(H0, order='C'), dtype=self._ctype),
            np.ascontiguousarray(np.ravel(H1, order='C'), dtype=self._ctype),
            dim, amps, bool(use_magnus), mode_sel))

    @_must_be_alive
    def equiprop(self, dt, *carr):
        """Propagator of the loaded Hamiltonians under the control waveforms ``*carr``.

        One 1D array of samples per control Hamiltonian (fewer arrays than Hamiltonians
        leave the rest at zero amplitude); a 2D array unpacks with ``*``.
        Returns a ``(dim, dim)`` ndarray (or a qutip object).
        """
        if self.amps < 0:
            raise RuntimeError("No hamiltonian set")
        amps = len(carr)
        if amps > self.amps:
            raise ValueError(f'Got {amps} amplitude arrays, but there are only {self.amps} Hamiltonians.')
        pts = np.shape(carr[0])[0]
        if any(np.shape(carri) != (pts,) for carri in carr):
            raise ValueError("All amplitude arrays must have the same length.")
        output = np.zeros(self.dim ** 2, dtype=self._ctype)
        carr = np.ascontiguousarray(np.ravel(np.asarray(carr, dtype=self._ctype), order='C'))
        self._check_error(self._call('Parament_equiprop', self._handle, carr, float(dt), int(pts), int(amps), output))
        output_data = np.ascontiguousarray(np.reshape(output, (self.dim, self.dim)))
        if self._use_qutip:
            output_data = qutip.Qobj(output_data, dims=self._qutip_H0.dims)
        return output_data

    # ---- additions (the C API has them; the reference's Python does not bind them) ----

    @_must_be_alive
    def set_iteration_cycles(self, cycles):
        """Force the Chebyshev order M (odd, >= 3); see Parament_setIterationCyclesManually."""
        self._check_error(self._call('Parament_setIterationCyclesManually', self._handle, int(cycles)))

    @_must_be_alive
    def automatic_iteration_cycles(self):
        """Restore the automatic choice of M from the iteration tables."""
        self._check_error(self._call('Parament_automaticIterationCycles', self._handle))

    @property
    def iteration_cycles(self):
        """Chebyshev order M used by the last successful :meth:`equiprop` (-1 before any)."""
        if self._handle is None:
            return -1
        return self._call('Parament_getIterationCycles', self._handle)

    @_must_be_alive
    def set_precision_mode(self, mode):
        """Arithmetic of a complex128 context: 'native' (default, FP64 throughout), 'mixed'
        (bf16 three-way split products), 'mixed_fp32' (FP32 products) or 'mixed_bf16x2'
        (bf16 hi+lo products); the mixed modes keep FP64 for the first-order terms and the
        upper product-tree levels."""
        if not self._use_doubles:
            raise ValueError("precision modes apply to 'fp64' contexts")
        code = {'native': PARAMENT_PRECISION_NATIVE, 'mixed': PARAMENT_PRECISION_MIXED,
                'mixed_fp32': PARAMENT_PRECISION_MIXED_FP32, 'mixed_bf16x2': PARAMENT_PRECISION_MIXED_BF16X2}.get(mode)
        if code is None:
            raise ValueError("mode must be 'native', 'mixed', 'mixed_fp32' or 'mixed_bf16x2'")
        self._check_error(self._lib.Parament_setPrecisionMode_fp64(self._handle, code))

    @_must_be_alive
    def set_kernel_path(self, path):
        """'auto' (fused on-chip kernels for dim <= 64) or 'stepwise' (batched step kernels)."""
        code = {'auto': 0, 'stepwise': 1}.get(path)
        if code is None:
            raise ValueError("path must be 'auto' or 'stepwise'")
        self._check_error(self._call('Parament_setKernelPath', self._handle, code))

    @_must_be_alive
    def equiprop_device(self, dt, carr_ptr, pts, amps, out_ptr):
        """Asynchronous equiprop on device pointers (controls and result stay in HBM)."""
        self._check_error(self._call('Parament_equipropDevice', self._handle, ctypes.c_void_p(carr_ptr), float(dt),
                                     int(pts), int(amps), ctypes.c_void_p(out_ptr)))

    @_must_be_alive
    def gate_fidelity_gradient(self, dt, *carr, target):
        """F = |Tr(target^H U)|^2 / dim^2 and dF/dc for every control sample (adjoint method).

        Real controls, any dim, every quadrature mode including Magnus.  Returns ``(F, grad)`` with ``grad`` shaped like
        the stacked controls, ``(amps, pts)``, in the context's precision.
        """
        if self.amps < 0:
            raise RuntimeError("No hamiltonian set")
        amps = len(carr)
        if amps > self.amps:
            raise ValueError(f'Got {amps} amplitude arrays, but there are only {self.amps} Hamiltonians.')
        pts = np.shape(carr[0])[0]
        if any(np.shape(carri) != (pts,) for carri in carr):
            raise ValueError("All amplitude arrays must have the same length.")
        c = np.asarray(carr)
        if np.iscomplexobj(c) and np.any(c.imag != 0):
            raise ValueError("gradients need real control amplitudes")
        ctype = np.complex128 if self._use_doubles else np.complex64
        rtype = np.float64 if self._use_doubles else np.float32
        c = np.ascontiguousarray(np.ravel(c.real.astype(ctype), order='C'))
        t = np.ascontiguousarray(np.ravel(np.asarray(target, dtype=ctype), order='C'))
        if t.size != self.dim ** 2:
            raise ValueError("target must be a dim x dim matrix")
        fid = ctypes.c_double() if self._use_doubles else ctypes.c_float()
        grad = np.zeros(amps * pts, dtype=rtype)
        self._check_error(self._call('Parament_gateFidelityGradient', self._handle, c, float(dt), int(pts), int(amps), t,
                                     ctypes.byref(fid), grad))
        return fid.value, grad.reshape(amps, pts)

    @_must_be_alive
    def gate_fidelity_gradient_device(self, dt, carr_ptr, pts, amps, target_ptr, fid_ptr, grad_ptr):
        """Device-pointer variant (asynchronous on the context's stream); reads the real parts of
        the complex control array at ``carr_ptr``."""
        self._check_error(self._call('Parament_gateFidelityGradientDevice', self._handle, ctypes.c_void_p(carr_ptr), float(dt),
                                     int(pts), int(amps), ctypes.c_void_p(target_ptr),
                                     ctypes.c_void_p(fid_ptr) if fid_ptr else None, ctypes.c_void_p(grad_ptr)))

    @_must_be_alive
    def set_gradient_checkpointing(self, segment=1, chunk=0):
        """Gradient memory policy: ``segment=0`` keeps the whole product tree (default of a new
        context); ``segment=1`` checkpoints at ~sqrt(N)-slice segments; larger values set the
        segment length.  ``chunk`` is the number of slices recomputed per pass (0: automatic)."""
        self._check_error(self._call('Parament_setGradientCheckpointing', self._handle, int(segment), int(chunk)))

    @property
    @_must_be_alive
    def gradient_workspace(self):
        """Device bytes the last gradient call worked in."""
        return int(self._call('Parament_getGradientWorkspace', self._handle))

    @_must_be_alive
    def set_stream(self, stream_ptr):
        """Run on the given HIP stream (an int handle, e.g. torch.cuda.current_stream().cuda_stream).

        0 / None is the device's null stream (PyTorch's default stream); passing back a value
        read from ``stream`` before restores the context's own stream."""
        self._check_error(self._call('Parament_setStream', self._handle, ctypes.c_void_p(stream_ptr or None)))

    @property
    def stream(self):
        return self._call('Parament_getStream', self._handle) if self._handle is not None else None

    @_must_be_alive
    def set_timing(self, enable=True):
        self._check_error(self._call('Parament_setTiming', self._handle, int(bool(enable))))

    @_must_be_alive
    def kernel_times(self):
        """(slice-kernel ms, product-tree ms) of the last equiprop, with timing enabled."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check_error(self._call('Parament_getKernelTimes', self._handle, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    @_must_be_alive
    def synchronize(self):
        self._check_error(self._call('Parament_synchronize', self._handle))

    @_must_be_alive
    def _get_error_message(self, code=None):
        if code is None:
            code = self._call('Parament_peekAtLastError', self._handle)
        return self._lib.Parament_errorMessage(code).decode()

    def _check_error(self, error_code):
        """Translate a C status code into the reference's exception types and messages."""
        if error_code == PARAMENT_STATUS_SUCCESS:
            return
        try:
            exception_class = _EXCEPTIONS[error_code]
        except KeyError as e:
            raise AssertionError('Unknown error code ') from e
        message = self._lib.Parament_errorMessage(error_code).decode()
        raise exception_class(f"Error code {error_code}: {message}")

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        self.destroy()
