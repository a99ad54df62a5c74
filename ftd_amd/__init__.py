"""ftd_amd -- MI355X-native X-engine HERK (V = X X^H, packed lower triangle).

The compute path is libftdx.so (hand-written HIP for gfx950 behind a C ABI,
include/ftdx.h); this package holds its build recipe and the Python host side.
"""
from . import ftdx  # noqa: F401
from .ftdx import (  # noqa: F401
    CutlassComplexGemm,
    ComputePrecision,
    DirectTileConfig,
    HerkMode,
    InputPrecision,
    OutputFormat,
    OutputPrecision,
    herk_int4,
    herkBatchedCutlassQC,
    herkBatchedCutlassQC_FP16,
    initCutlassApi,
)
