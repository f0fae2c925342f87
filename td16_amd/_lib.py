"""ctypes binding of libtd16.so (the C ABI declared in include/td16_abi.h).

The library is built in-tree (``td16_amd/libtd16.so``) by
``__graft_entry__.build()`` or ``make -C td16_amd/csrc``.  Loading fails
loudly when it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtd16.so")

MAX_N = 17
MAX_DEPTH = 9
CERT_KEY_BYTES = 11
CERT_VALUE_BYTES = 17
BRACELET_PAIRS, BRACELET_WORDS = 0, 1
NOALLOC_MAX_REQUESTS = 32768
NOALLOC_QUEUE_ITEMS = 32768

OK = 0
INVALID_ARGUMENT = 1
OUT_OF_MEMORY = 2
DEVICE_ERROR = 3
NOT_INITIALIZED = 4
QUEUE_FULL = 5

MOVES_REF, MOVES_ALL = 0, 1
FIRST, COUNT_ALL = 0, 1

_STATUS_NAMES = {
    OK: "OK",
    INVALID_ARGUMENT: "INVALID_ARGUMENT",
    OUT_OF_MEMORY: "OUT_OF_MEMORY",
    DEVICE_ERROR: "DEVICE_ERROR",
    NOT_INITIALIZED: "NOT_INITIALIZED",
    QUEUE_FULL: "QUEUE_FULL",
}


class Td16Error(RuntimeError):
    """A non-OK td16_status; the reference raises java.lang.RuntimeException."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status
        self.status_name = _STATUS_NAMES.get(status, str(status))


class GpuMemoryExhausted(Td16Error):
    """TD16_OUT_OF_MEMORY (the reference's GpuMemoryExhaustedException)."""


class Request(ctypes.Structure):
    _fields_ = [
        ("beta", ctypes.c_uint8 * MAX_N),
        ("omega", ctypes.c_uint8 * MAX_N),
        ("lex_order", ctypes.c_uint8),
        ("pad0", ctypes.c_uint8),
        ("n", ctypes.c_int32),
        ("initial_even", ctypes.c_int32),
        ("min_rate", ctypes.c_float),
        ("max_moves", ctypes.c_int32),
        ("full_sorting", ctypes.c_int32),
        ("moveset", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("iddfs", ctypes.c_int32),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("found", ctypes.c_int32),
        ("depth", ctypes.c_int32),
        ("split", ctypes.c_int32),
        ("moves", (ctypes.c_int32 * 3) * MAX_DEPTH),
        ("expanded", ctypes.c_uint64),
        ("solutions", ctypes.c_uint64),
        ("device_us", ctypes.c_double),
        ("latency_us", ctypes.c_double),
        ("queue_us", ctypes.c_double),
    ]

    def witness(self):
        return [tuple(self.moves[d]) for d in range(self.depth if self.found else 0)]


class BatchOptions(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("waves_per_cu", ctypes.c_int32),
        ("split_depth", ctypes.c_int32),
        ("donate_max", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("repeats", ctypes.c_int32),
        ("queue_bytes", ctypes.c_int64),
        ("check_every", ctypes.c_int32),
        ("host_threads", ctypes.c_int32),
        ("count_candidates", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("kernel_ms", ctypes.c_double),
        ("wall_ms", ctypes.c_double),
        ("expanded", ctypes.c_uint64),
        ("solutions", ctypes.c_uint64),
        ("donations", ctypes.c_uint64),
        ("waves", ctypes.c_int32),
        ("blocks", ctypes.c_int32),
        ("prepare_ms", ctypes.c_double),
        ("upload_ms", ctypes.c_double),
        ("decode_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("candidates", ctypes.c_uint64),
    ]


class ServeOptions(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("slots", ctypes.c_int32),
        ("queue_capacity", ctypes.c_int32),
        ("waves_per_cu", ctypes.c_int32),
        ("slot_bytes", ctypes.c_int64),
        ("deadline_ms", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class ServeStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("rejected", ctypes.c_uint64),
        ("p50_us", ctypes.c_double),
        ("p99_us", ctypes.c_double),
        ("mean_us", ctypes.c_double),
        ("in_flight", ctypes.c_int32),
        ("waiting", ctypes.c_int32),
    ]


# name -> (restype, argtypes); every td16_* symbol declared in include/td16_abi.h
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
class TunePoint(ctypes.Structure):
    _fields_ = [
        ("waves_per_cu", ctypes.c_int32),
        ("split_depth", ctypes.c_int32),
        ("donate_max", ctypes.c_int32),
        ("check_every", ctypes.c_int32),
        ("waves", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("kernel_ms", ctypes.c_double),
    ]


_c = ctypes.c_char_p
_i = ctypes.c_int
SIGNATURES = {
    "td16_init_slots": (_i, [_i, _i32p, _i, ctypes.c_int64, _c, _i]),
    "td16_shutdown": (_i, [_c, _i]),
    "td16_slot_count": (_i, []),
    "td16_slot_device": (_i, [_i]),
    "td16_search": (_i, [_i, _u8p, _u8p, _i, _i, ctypes.c_float, _i, _i, _i32p, _i32p, _c, _i]),
    "td16_search_ex": (_i, [_i, ctypes.POINTER(Request), ctypes.POINTER(Result), _c, _i]),
    "td16_batch": (_i, [ctypes.POINTER(Request), _i, ctypes.POINTER(Result),
                        ctypes.POINTER(BatchOptions), ctypes.POINTER(BatchStats), _c, _i]),
    "td16_batch_stream": (_i, [ctypes.POINTER(Request), _i, _i, ctypes.POINTER(Result),
                               ctypes.POINTER(BatchOptions), ctypes.POINTER(BatchStats), _c, _i]),
    "td16_batch_noalloc": (_i, [ctypes.POINTER(Request), _i, ctypes.POINTER(Result),
                                ctypes.POINTER(BatchOptions), ctypes.POINTER(BatchStats), _c, _i]),
    "td16_autotune": (_i, [ctypes.POINTER(Request), _i, ctypes.POINTER(Result),
                           ctypes.POINTER(BatchOptions), _i, ctypes.POINTER(BatchOptions),
                           ctypes.POINTER(TunePoint), _i, _i32p, _c, _i]),
    "td16_serve_start": (_i, [ctypes.POINTER(ServeOptions), _c, _i]),
    "td16_serve_submit": (_i, [ctypes.POINTER(Request), _i, _u64p, _c, _i]),
    "td16_serve_wait": (_i, [ctypes.c_uint64, ctypes.POINTER(Result), _c, _i]),
    "td16_serve_search": (_i, [ctypes.POINTER(Request), ctypes.POINTER(Result), _c, _i]),
    "td16_serve_stats_get": (_i, [ctypes.POINTER(ServeStats), _i, _c, _i]),
    "td16_serve_stop": (_i, [_c, _i]),
    "td16_serve_closed_loop": (_i, [ctypes.POINTER(Request), _i, _i, _i, ctypes.POINTER(Result),
                                    ctypes.POINTER(ctypes.c_double), _c, _i]),
    "td16_serve_replay": (_i, [ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_double), _i,
                               ctypes.POINTER(Result), ctypes.POINTER(ctypes.c_double), _c, _i]),
    "td16_probe_valu_peak": (_i, [_i, _i, ctypes.POINTER(ctypes.c_double), _c, _i]),
    "td16_probe_copy": (_i, [_i, ctypes.c_int64, ctypes.POINTER(ctypes.c_double), _c, _i]),
    "td16_positions_to_symbols": (_i, [_u8p, _i, _i32p, _i, _i32p, _c, _i]),
    "td16_validate": (_i, [_u8p, _u8p, _i, _i32p, _i, _i, _c, _i]),
    "td16_codec_key": (_i, [_u8p, _i, _u8p, _i, _i32p, _c, _i]),
    "td16_codec_value": (_i, [_i32p, _i, _u8p, _i, _i32p, _c, _i]),
    "td16_codec_decode_key": (_i, [_u8p, _i, _i, _u8p, _c, _i]),
    "td16_codec_decode_value": (_i, [_u8p, _i, _i32p, _i, _i32p, _c, _i]),
    "td16_certify_batch": (_i, [ctypes.POINTER(Request), ctypes.POINTER(Result), _i, _i, _u8p,
                                _u8p, _i32p, _i32p, _c, _i]),
    "td16_check_batch": (_i, [_u8p, _u8p, _i32p, _i32p, _i, _u8p, _i32p, _i, _i32p, _c, _i]),
    "td16_bracelets": (_i, [_i32p, _u8p, _i, _i, _i, _u8p, ctypes.c_int64,
                            ctypes.POINTER(ctypes.c_int64), _c, _i]),
    "td16_bracelet_canonical": (_i, [_u8p, _u8p, _i, _i32p, _u8p, _i32p, _u8p, _u8p, _c, _i]),
    "td16_gen_uniform": (_i, [ctypes.c_uint64, _i, _i, _u64p]),
    "td16_gen_products": (_i, [ctypes.c_uint64, _i, _i, _i, _i, _i, _u8p, _u8p, _u64p]),
    "td16_gen_exponential": (_i, [ctypes.c_uint64, _i, ctypes.POINTER(ctypes.c_double)]),
    "td16_linear_request": (_i, [ctypes.c_uint64, _i, _i, _i, _i, _i, ctypes.POINTER(Request)]),
    "td16_version": (ctypes.c_char_p, []),
}

_lib = None


def lib():
    """The loaded library; raises when it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "or `make -C td16_amd/csrc` (there is no fallback implementation)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, msg: ctypes.Array) -> None:
    if status != OK:
        text = msg.value.decode("utf-8", "replace") if msg is not None else ""
        if status == OUT_OF_MEMORY:
            raise GpuMemoryExhausted(status, text)
        raise Td16Error(status, text)


def msgbuf():
    return ctypes.create_string_buffer(512)
