"""MI355X implementation of mdSpherocyl's per-step force path.

HIP kernels for gfx950 behind the C ABI of ``include/msc.h``
(``mdspherocyl_amd/libmsc_hip.so``); ``MdSystem`` mirrors the reference
routines over that ABI and ``configs`` builds the benchmark systems.
"""
from ._lib import (EXPORTS, MSC_ALL, MSC_FP64, MSC_MIXED, OBS_NAMES, MscError, ReplicaDesc,  # noqa: F401
                   LIBPATH, load)
from .system import MdSystem, Overlap, ini_state, mt19937_state, replica_desc  # noqa: F401
