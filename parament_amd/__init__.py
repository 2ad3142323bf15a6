"""MI355X-native equiprop: drop-in for pyparament (``import parament_amd as parament``).

Exports the reference package's names (src/python/pyparament/parament/__init__.py:15-16):
``Parament``, ``expm``, ``device_info``, plus the C constants.  Loading fails with an
ImportError when libparament.so has not been built; there is no CPU fallback.
"""
from .constants import *  # noqa: F401,F403
from .parament import Parament
from .debug_functions import expm, device_info
from . import debug_functions, parament, paramentlib  # noqa: F401

__all__ = ["Parament", "expm", "device_info"]
