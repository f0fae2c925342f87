"""td16_amd: MI355X-native witness search for the td16 emitter.

The product is libtd16.so (hand-written HIP for gfx950 behind the C ABI in
include/td16_abi.h).  This package binds it with ctypes and mirrors the
reference's Java-side contract around it: the bounded slot dispatcher
(GPUSortingSearch) and the witness facade (search -> replay -> certificate
sink).
"""
from ._lib import (COUNT_ALL, DEVICE_ERROR, FIRST, GpuMemoryExhausted, INVALID_ARGUMENT,
                   MOVES_ALL, MOVES_REF, NOT_INITIALIZED, OK, OUT_OF_MEMORY, QUEUE_FULL,
                   Td16Error, LIB_PATH)
from . import native
from .gpu_sorting_search import GPUSortingSearch
from .witness import MemoryCertificateSink, WitnessSearch

__all__ = [
    "native", "GPUSortingSearch", "WitnessSearch",
    "MemoryCertificateSink", "Td16Error", "GpuMemoryExhausted", "MOVES_REF", "MOVES_ALL",
    "FIRST", "COUNT_ALL", "OK", "INVALID_ARGUMENT", "OUT_OF_MEMORY", "DEVICE_ERROR",
    "NOT_INITIALIZED", "QUEUE_FULL", "LIB_PATH",
]
