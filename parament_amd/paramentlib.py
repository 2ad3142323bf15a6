"""Locate and load libparament.so and declare its C signatures
(reference: src/python/pyparament/parament/paramentlib.py:16-72).

Same environment variables as the reference: PARAMENT_LIB_DIR names a directory holding
libparament.so; USE_SHARED_PARAMENT searches the system library path; otherwise the
library built next to this file is used.  A missing library raises at import, so no
code path can run without the HIP implementation.
"""
import ctypes
import os
import pathlib

import numpy as np

USE_SHARED_PARAMENT = os.environ.get("USE_SHARED_PARAMENT")
PARAMENT_LIB_DIR = os.environ.get("PARAMENT_LIB_DIR")

if os.name != "posix":
    raise RuntimeError("Don't know how to load library on " + os.name)

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

lib.Parament_setPrecisionMode_fp64.argtypes = [c_ParamentContext_p, ctypes.c_int]
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
