"""Locate and load libparament.so and declare its C signatures
(reference: src/python/pyparament/parament/paramentlib.py:16-72).

Same environment variables as the reference: PARAMENT_LIB_DIR names a directory holding
libparament.so; USE_SHARED_PARAMENT searches the system library path; otherwise the
library built next to this file is used.  A missing library raises at import, so no
code path can run without the HIP implementation.
"""
import ctypes
import importlib.util
import os
import pathlib

import numpy as np

USE_SHARED_PARAMENT = os.environ.get("USE_SHARED_PARAMENT")
PARAMENT_LIB_DIR = os.environ.get("PARAMENT_LIB_DIR")

if os.name != "posix":
    raise RuntimeError("Don't know how to load library on " + os.name)



def _share_torch_hip_runtime():
    """Load PyTorch-ROCm's HIP runtime (if installed) before libparament.so.

    torch ships its own libamdhip64 with the same SONAME as the system one; whichever is
    loaded first is then used by both, so torch tensors, torch streams and this library
    share one runtime in the process.  Loading the system runtime first would leave two
    runtimes competing for the device.  Set PARAMENT_OWN_HIP_RUNTIME=1 to skip this.
    """
    if os.environ.get("PARAMENT_OWN_HIP_RUNTIME"):
        return
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return
    cand = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(cand):
        ctypes.CDLL(cand, mode=ctypes.RTLD_GLOBAL)


_share_torch_hip_runtime()

if PARAMENT_LIB_DIR:
    lib_path = str(pathlib.Path(PARAMENT_LIB_DIR) / "libparament.so")
elif not USE_SHARED_PARAMENT:
    lib_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libparament.so")
else:
    lib_path = "libparament.so"

if not USE_SHARED_PARAMENT and not os.path.exists(lib_path):
    raise ImportError(f"libparament.so not found at {lib_path}; build it with "
                      "`python -c 'import __graft_entry__ as g; g.build()'` or parament_amd/build.py")

lib = ctypes.CDLL(lib_path)

c_ParamentContext_p = ctypes.c_void_p
c_cuComplex_p = np.ctypeslib.ndpointer(np.complex64, flags="C_CONTIGUOUS")
c_cuDoubleComplex_p = np.ctypeslib.ndpointer(np.complex128, flags="C_CONTIGUOUS")

for sfx, cptr in (("", c_cuComplex_p), ("_fp64", c_cuDoubleComplex_p)):
    getattr(lib, "Parament_create" + sfx).argtypes = [ctypes.POINTER(c_ParamentContext_p)]
    getattr(lib, "Parament_destroy" + sfx).argtypes = [c_ParamentContext_p]
    getattr(lib, "Parament_setHamiltonian" + sfx).argtypes = [c_ParamentContext_p, cptr, cptr, ctypes.c_uint,
                                                              ctypes.c_uint, ctypes.c_bool, ctypes.c_int]
    getattr(lib, "Parament_equiprop" + sfx).argtypes = [c_ParamentContext_p, cptr, ctypes.c_double, ctypes.c_uint,
                                                        ctypes.c_uint, cptr]
    getattr(lib, "Parament_setIterationCyclesManually" + sfx).argtypes = [c_ParamentContext_p, ctypes.c_uint]
    getattr(lib, "Parament_automaticIterationCycles" + sfx).argtypes = [c_ParamentContext_p]
    getattr(lib, "Parament_peekAtLastError" + sfx).argtypes = [c_ParamentContext_p]
    getattr(lib, "Parament_getIterationCycles" + sfx).argtypes = [c_ParamentContext_p]
    getattr(lib, "Parament_getIterationCycles" + sfx).restype = ctypes.c_int
    getattr(lib, "Parament_equipropDevice" + sfx).argtypes = [c_ParamentContext_p, ctypes.c_void_p, ctypes.c_double,
                                                              ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p]
    getattr(lib, "Parament_synchronize" + sfx).argtypes = [c_ParamentContext_p]
    getattr(lib, "Parament_setKernelPath" + sfx).argtypes = [c_ParamentContext_p, ctypes.c_int]
    getattr(lib, "Parament_setStream" + sfx).argtypes = [c_ParamentContext_p, ctypes.c_void_p]
    getattr(lib, "Parament_getStream" + sfx).argtypes = [c_ParamentContext_p]
    getattr(lib, "Parament_getStream" + sfx).restype = ctypes.c_void_p
    getattr(lib, "Parament_setTiming" + sfx).argtypes = [c_ParamentContext_p, ctypes.c_int]
    getattr(lib, "Parament_getKernelTimes" + sfx).argtypes = [c_ParamentContext_p, ctypes.POINTER(ctypes.c_double),
                                                              ctypes.POINTER(ctypes.c_double)]

for sfx, cptr, real in (("", c_cuComplex_p, np.float32), ("_fp64", c_cuDoubleComplex_p, np.float64)):
    rptr = ctypes.POINTER(ctypes.c_float if real is np.float32 else ctypes.c_double)
    getattr(lib, "Parament_gateFidelityGradient" + sfx).argtypes = [
        c_ParamentContext_p, cptr, ctypes.c_double, ctypes.c_uint, ctypes.c_uint, cptr, rptr,
        np.ctypeslib.ndpointer(real, flags="C_CONTIGUOUS")]
    getattr(lib, "Parament_gateFidelityGradientDevice" + sfx).argtypes = [
        c_ParamentContext_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p,
        ctypes.c_void_p, ctypes.c_void_p]
    getattr(lib, "Parament_setGradientCheckpointing" + sfx).argtypes = [c_ParamentContext_p, ctypes.c_uint,
                                                                        ctypes.c_uint]
    getattr(lib, "Parament_getGradientWorkspace" + sfx).argtypes = [c_ParamentContext_p]
    getattr(lib, "Parament_getGradientWorkspace" + sfx).restype = ctypes.c_size_t
for sfx in ("", "_fp64"):
    getattr(lib, "Parament_setPrecisionMode" + sfx).argtypes = [c_ParamentContext_p, ctypes.c_int]
    getattr(lib, "Parament_getPrecisionMode" + sfx).argtypes = [c_ParamentContext_p]
    getattr(lib, "Parament_getPrecisionMode" + sfx).restype = ctypes.c_int
_uint_p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
lib.Parament_ensembleCreate.argtypes = [ctypes.POINTER(c_ParamentContext_p), ctypes.c_uint, _uint_p, _uint_p, _uint_p,
                                        np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS"), ctypes.c_int]
lib.Parament_ensembleDestroy.argtypes = [c_ParamentContext_p]
lib.Parament_ensembleSetHamiltonians.argtypes = [c_ParamentContext_p, c_cuComplex_p]
lib.Parament_ensembleSetExecution.argtypes = [c_ParamentContext_p, ctypes.c_int]
lib.Parament_ensembleTune.argtypes = [c_ParamentContext_p]
lib.Parament_ensembleGetTuning.argtypes = [c_ParamentContext_p, np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")]
lib.Parament_ensembleEquiprop.argtypes = [c_ParamentContext_p, c_cuComplex_p, c_cuComplex_p]
lib.Parament_ensembleEquipropDevice.argtypes = [c_ParamentContext_p, ctypes.c_void_p, ctypes.c_void_p]
lib.Parament_ensembleSynchronize.argtypes = [c_ParamentContext_p]
lib.Parament_ensemblePeekAtLastError.argtypes = [c_ParamentContext_p]
lib.Parament_ensembleGateFidelityGradient.argtypes = [c_ParamentContext_p, c_cuComplex_p, c_cuComplex_p,
                                                      np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS"),
                                                      np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")]
lib.Parament_ensembleGateFidelityGradientDevice.argtypes = [c_ParamentContext_p, ctypes.c_void_p, ctypes.c_void_p,
                                                            ctypes.c_void_p, ctypes.c_void_p]
lib.Parament_equipropDeviceNoAlloc_fp64.argtypes = [c_ParamentContext_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_uint,
                                                    ctypes.c_uint, ctypes.c_void_p]
lib.Parament_errorMessage.argtypes = [ctypes.c_int]
lib.Parament_errorMessage.restype = ctypes.c_char_p
lib.Parament_selectIterationCycles_fp32.argtypes = [ctypes.c_double, ctypes.c_double]
lib.Parament_selectIterationCycles_fp32.restype = ctypes.c_int
lib.Parament_selectIterationCycles_fp64.argtypes = [ctypes.c_double, ctypes.c_double]
lib.Parament_selectIterationCycles_fp64.restype = ctypes.c_int
lib.OneNorm.argtypes = [c_cuComplex_p, ctypes.c_uint]
lib.OneNorm.restype = ctypes.c_double
lib.OneNorm_fp64.argtypes = [c_cuDoubleComplex_p, ctypes.c_uint]
lib.OneNorm_fp64.restype = ctypes.c_double
lib.device_info.argtypes = []
lib.device_info.restype = None
