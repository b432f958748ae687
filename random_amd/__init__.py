"""MI355X-native Kirchhoff prestack time migration.

The hot path of ``Kirchhoff.c`` (obriengareth/Random) -- ``kirchhoff(job_id)``
-- rebuilt as a C host layer plus hand-written gfx950 HIP kernels in
``libktm.so``, exposed through the C ABI in ``include/ktm.h``.  This package
is the Python side of that boundary: ctypes bindings, the configuration
mirror of the reference's macros, device-resident plans driven with torch
tensors, seeded synthetic surveys, the adjoint pair and least-squares
migration.
"""
from ._lib import (ACCUM_KAHAN, ACCUM_PLAIN, INDEX_FAST, PREP_CLOSED, PREP_REFERENCE,
                   TRACE_BF16, TRACE_F16, TRACE_F32, KtmError)
from .ktm import Config, Plan, migrate_host, setup_info

__all__ = [
    "ACCUM_KAHAN", "ACCUM_PLAIN", "INDEX_FAST", "PREP_CLOSED", "PREP_REFERENCE", "TRACE_BF16", "TRACE_F16",
    "TRACE_F32", "KtmError", "Config", "Plan", "migrate_host", "setup_info",
]
