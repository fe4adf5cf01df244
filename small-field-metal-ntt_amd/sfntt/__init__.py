"""MI355X (gfx950) small-field NTT / coset-LDE engine: Python front-end.

All compute runs in libsfntt.so (hand-written HIP kernels behind a C ABI,
include/sfntt.h); this package only marshals arguments.
"""
from . import _lib
from ._lib import (CT_DIT_R2, CT_GS_R2, CT_GS_R4, FIELD_BABYBEAR, FIELD_KOALABEAR, FIELD_M31, FORWARD,
                   INVERSE, LAYOUT_ROW_MAJOR, LAYOUT_SOA, STOCKHAM_R2, VARIANT_AUTO)
from .api import *  # noqa: F401,F403
from .api import (BB_VARIANTS, KB_VARIANTS, M31_VARIANTS, BbBatchNtt, CosetLdeBatch, DeviceBuffer,
                  HipContext, KbBatchNtt, NttError, default_context)

__all__ = ["_lib", "HipContext", "DeviceBuffer", "default_context", "NttError", "CosetLdeBatch",
           "BbBatchNtt", "KbBatchNtt", "M31_VARIANTS", "BB_VARIANTS", "KB_VARIANTS", "FIELD_M31",
           "FIELD_BABYBEAR", "FIELD_KOALABEAR", "FORWARD",
           "INVERSE", "LAYOUT_SOA", "LAYOUT_ROW_MAJOR", "CT_DIT_R2", "CT_GS_R2", "STOCKHAM_R2",
           "CT_GS_R4", "VARIANT_AUTO"]
