"""Pythonic wrappers over the C ABI of libtd16.so.

Requests use the reference's form: ``beta`` is the cyclic sequence and
``omega`` the one-line algebraic permutation, both permutations of 0..n-1
(reference: emitter/src/main/java/br/unb/cic/tdp/base/GPUSortingSearch.java:
117, 203-240).  Results carry position triples exactly as the reference's
native search returns them (emitter/src/native/jni/jni_sorter.cu:100-110).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib as _lib_mod
from ._lib import (FIRST, COUNT_ALL, MOVES_ALL, MOVES_REF, BatchOptions, BatchStats,
                   Request, Result, ServeOptions, ServeStats, Td16Error, TunePoint, check, lib,
                   msgbuf)

Move = Tuple[int, int, int]


def _u8(seq: Sequence[int]):
    return (ctypes.c_uint8 * max(1, len(seq)))(*seq)


def make_request(beta: Sequence[int], omega: Sequence[int], initial_even: int, min_rate: float,
                 max_moves: int, full_sorting: bool = True, moveset: int = MOVES_REF,
                 mode: int = FIRST, iddfs: bool = False, lex_order: bool = False) -> Request:
    if len(beta) != len(omega):
        raise Td16Error(1, "beta and omega must have equal lengths")
    if len(beta) > 17:
        raise Td16Error(1, "performGpuSortingSearch: n exceeds MAX_N")
    r = Request()
    for q, x in enumerate(beta):
        r.beta[q] = x & 0xFF
    for q, x in enumerate(omega):
        r.omega[q] = x & 0xFF
    r.n = len(beta)
    r.lex_order = 1 if lex_order else 0
    r.initial_even = int(initial_even)
    r.min_rate = float(min_rate)
    r.max_moves = int(max_moves)
    r.full_sorting = 1 if full_sorting else 0
    r.moveset = int(moveset)
    r.mode = int(mode)
    r.iddfs = 1 if iddfs else 0
    return r


def linear_request(perm: int, n: int, k: int, moveset: int = MOVES_REF, mode: int = FIRST,
                   iddfs: bool = False, lex_order: bool = False) -> Request:
    """A nibble-packed linear permutation as a request: beta = [pi, n],
    omega(y) = pred(y) + 1, min_rate by the emitter's convention."""
    r = Request()
    st = lib().td16_linear_request(int(perm), n, k, moveset, mode, 1 if iddfs else 0,
                                   ctypes.byref(r))
    if st != 0:
        raise Td16Error(st, "invalid linear instance")
    r.lex_order = 1 if lex_order else 0
    return r


# ---------------------------------------------------------------------------
# slots and the drop-in search
# ---------------------------------------------------------------------------

def init_slots(total_slots: int, devices: Sequence[int], queue_bytes_budget: int) -> None:
    arr = (ctypes.c_int32 * max(1, len(devices)))(*devices)
    m = msgbuf()
    check(lib().td16_init_slots(total_slots, arr if devices else None, len(devices),
                                int(queue_bytes_budget), m, len(m)), m)


def shutdown() -> None:
    m = msgbuf()
    check(lib().td16_shutdown(m, len(m)), m)


def slot_count() -> int:
    return lib().td16_slot_count()


def slot_device(slot: int) -> int:
    return lib().td16_slot_device(slot)


def search_positions(slot: int, beta: Sequence[int], omega: Sequence[int], initial_even: int,
                     min_rate: float, max_moves: int, full_sorting: bool = True) -> List[Move]:
    """The drop-in: jni_sorter.cu:49-117 -> int[len][3] positions."""
    if beta is None or omega is None:
        raise Td16Error(1, "beta and omega must not be null")
    if len(beta) != len(omega):
        raise Td16Error(1, "beta and omega must have equal lengths")
    out = (ctypes.c_int32 * 27)()
    ln = ctypes.c_int32(0)
    m = msgbuf()
    check(lib().td16_search(slot, _u8([x & 0xFF for x in beta]), _u8([x & 0xFF for x in omega]),
                            len(beta), int(initial_even), float(min_rate), int(max_moves),
                            1 if full_sorting else 0, out, ctypes.byref(ln), m, len(m)), m)
    return [(out[3 * d], out[3 * d + 1], out[3 * d + 2]) for d in range(ln.value)]


def search_ex(slot: int, req: Request) -> Result:
    res = Result()
    m = msgbuf()
    check(lib().td16_search_ex(slot, ctypes.byref(req), ctypes.byref(res), m, len(m)), m)
    return res


# ---------------------------------------------------------------------------
# batched
# ---------------------------------------------------------------------------

@dataclass
class BatchRun:
    results: ctypes.Array
    stats: BatchStats


def batch(reqs: Sequence[Request], device: int = 0, waves_per_cu: int = 0, split_depth: int = 0,
          donate_max: int = 0, use_graph: bool = False, repeats: int = 1,
          queue_bytes: int = 0, check_every: int = 0, host_threads: int = 0,
          out: Optional[ctypes.Array] = None, count_candidates: bool = False) -> BatchRun:
    """One persistent launch over ``reqs`` (``td16_batch``).  ``out``: a
    result array of at least len(reqs) entries to fill (a new one
    otherwise)."""
    n = len(reqs)
    arr = (Request * max(1, n))(*reqs) if not isinstance(reqs, ctypes.Array) else reqs
    if out is None or len(out) < n:
        out = (Result * max(1, n))()
    opt = BatchOptions(device, waves_per_cu, split_depth, donate_max, 1 if use_graph else 0,
                       repeats, queue_bytes, check_every, host_threads,
                       1 if count_candidates else 0, 0)
    st = BatchStats()
    m = msgbuf()
    check(lib().td16_batch(arr, n, out, ctypes.byref(opt), ctypes.byref(st), m, len(m)), m)
    return BatchRun(out, st)


def batch_stream(reqs: Sequence[Request], chunk: int, device: int = 0, waves_per_cu: int = 0,
                 split_depth: int = 0, donate_max: int = 0, queue_bytes: int = 0,
                 check_every: int = 0, host_threads: int = 0,
                 out: Optional[ctypes.Array] = None) -> BatchRun:
    """Consecutive launches of at most ``chunk`` requests with the host's
    preparation and decoding overlapped with the searches
    (``td16_batch_stream``); ``out`` as in :func:`batch`."""
    n = len(reqs)
    arr = (Request * max(1, n))(*reqs) if not isinstance(reqs, ctypes.Array) else reqs
    if out is None or len(out) < n:
        out = (Result * max(1, n))()
    opt = BatchOptions(device, waves_per_cu, split_depth, donate_max, 0, 1, queue_bytes,
                       check_every, host_threads)
    st = BatchStats()
    m = msgbuf()
    check(lib().td16_batch_stream(arr, n, chunk, out, ctypes.byref(opt), ctypes.byref(st), m,
                                  len(m)), m)
    return BatchRun(out, st)


def batch_noalloc(reqs: Sequence[Request], device: int = 0, waves_per_cu: int = 0,
                  split_depth: int = 0, donate_max: int = 0, repeats: int = 1,
                  check_every: int = 0, use_graph: bool = False) -> BatchRun:
    """:func:`batch` through ``td16_batch_noalloc``: no device allocation, at
    most ``NOALLOC_MAX_REQUESTS`` requests per call."""
    n = len(reqs)
    arr = (Request * max(1, n))(*reqs) if not isinstance(reqs, ctypes.Array) else reqs
    out = (Result * max(1, n))()
    opt = BatchOptions(device, waves_per_cu, split_depth, donate_max, 1 if use_graph else 0,
                       repeats, 0, check_every, 0)
    st = BatchStats()
    m = msgbuf()
    check(lib().td16_batch_noalloc(arr, n, out, ctypes.byref(opt), ctypes.byref(st), m, len(m)),
          m)
    return BatchRun(out, st)


@dataclass
class Tuning:
    best: dict            # waves_per_cu, split_depth, donate_max, check_every
    points: List[dict]    # every timed configuration, in the order tried
    results: ctypes.Array


def autotune(reqs: Sequence[Request], device: int = 0, repeats: int = 3, use_graph: bool = True,
             queue_bytes: int = 0, waves_per_cu: int = 0, split_depth: int = 0,
             donate_max: int = 0, check_every: int = 0, max_points: int = 64) -> Tuning:
    """Times the batch over the knob grid of ``td16_autotune`` and returns the
    fastest setting; pass ``best`` to :func:`batch`.  Raises ``Td16Error`` if
    any setting changed an answer."""
    n = len(reqs)
    arr = (Request * max(1, n))(*reqs) if not isinstance(reqs, ctypes.Array) else reqs
    out = (Result * max(1, n))()
    base = BatchOptions(device, waves_per_cu, split_depth, donate_max, 1 if use_graph else 0, 1,
                        queue_bytes, check_every, 0)
    best = BatchOptions()
    pts = (TunePoint * max_points)()
    npts = ctypes.c_int32(0)
    m = msgbuf()
    check(lib().td16_autotune(arr, n, out, ctypes.byref(base), repeats, ctypes.byref(best), pts,
                              max_points, ctypes.byref(npts), m, len(m)), m)
    points = [dict(waves_per_cu=p.waves_per_cu, split_depth=p.split_depth,
                   donate_max=p.donate_max, check_every=p.check_every, waves=p.waves,
                   kernel_ms=p.kernel_ms)
              for p in pts[:min(npts.value, max_points)]]
    return Tuning(dict(waves_per_cu=best.waves_per_cu, split_depth=best.split_depth,
                       donate_max=best.donate_max, check_every=best.check_every), points, out)


# ---------------------------------------------------------------------------
# served loop
# ---------------------------------------------------------------------------

def serve_start(device: int = 0, slots: int = 64, queue_capacity: int = 512,
                slot_bytes: int = 1 << 30, waves_per_cu: int = 0, deadline_ms: int = 0) -> None:
    opt = ServeOptions(device, slots, queue_capacity, waves_per_cu, slot_bytes, deadline_ms, 0)
    m = msgbuf()
    check(lib().td16_serve_start(ctypes.byref(opt), m, len(m)), m)


def serve_submit(req: Request, nonblocking: bool = False) -> int:
    t = ctypes.c_uint64(0)
    m = msgbuf()
    check(lib().td16_serve_submit(ctypes.byref(req), 1 if nonblocking else 0, ctypes.byref(t),
                                  m, len(m)), m)
    return t.value


def serve_wait(ticket: int) -> Result:
    res = Result()
    m = msgbuf()
    check(lib().td16_serve_wait(ticket, ctypes.byref(res), m, len(m)), m)
    return res


def serve_search(req: Request) -> Result:
    res = Result()
    m = msgbuf()
    check(lib().td16_serve_search(ctypes.byref(req), ctypes.byref(res), m, len(m)), m)
    return res


def serve_stats(reset_latency: bool = False) -> ServeStats:
    st = ServeStats()
    m = msgbuf()
    check(lib().td16_serve_stats_get(ctypes.byref(st), 1 if reset_latency else 0, m, len(m)), m)
    return st


def serve_replay(reqs: Sequence[Request], arrival_us: Sequence[float]):
    """Open-loop load: request q arrives arrival_us[q] after the start.
    Returns (results, elapsed_us, stats) with p50/p99 from scheduled arrival."""
    n = len(reqs)
    arr = (Request * max(1, n))(*reqs)
    arrivals = (ctypes.c_double * max(1, n))(*arrival_us)
    out = (Result * max(1, n))()
    el = ctypes.c_double(0)
    m = msgbuf()
    check(lib().td16_serve_replay(arr, arrivals, n, out, ctypes.byref(el), m, len(m)), m)
    return out, el.value, serve_stats()


def serve_closed_loop(reqs: Sequence[Request], producers: int, window: int = 1):
    """Closed-loop load: `producers` threads each keep up to `window`
    requests in flight, with no pacing.  Returns (results, elapsed_us,
    stats)."""
    n = len(reqs)
    arr = (Request * max(1, n))(*reqs) if not isinstance(reqs, ctypes.Array) else reqs
    out = (Result * max(1, n))()
    el = ctypes.c_double(0)
    m = msgbuf()
    check(lib().td16_serve_closed_loop(arr, n, producers, window, out, ctypes.byref(el), m,
                                       len(m)), m)
    return out, el.value, serve_stats()


def probe_valu_peak(device: int = 0, iterations: int = 20000) -> float:
    """Sustained int32 VALU rate (10^12 lane-ops/s)."""
    v = ctypes.c_double(0)
    m = msgbuf()
    check(lib().td16_probe_valu_peak(device, iterations, ctypes.byref(v), m, len(m)), m)
    return v.value


def probe_copy(device: int = 0, nbytes: int = 256 << 20) -> float:
    """Read+write GB/s of a plain device copy of nbytes."""
    v = ctypes.c_double(0)
    m = msgbuf()
    check(lib().td16_probe_copy(device, int(nbytes), ctypes.byref(v), m, len(m)), m)
    return v.value


def serve_stop() -> None:
    m = msgbuf()
    check(lib().td16_serve_stop(m, len(m)), m)


# ---------------------------------------------------------------------------
# replay and certificate format
# ---------------------------------------------------------------------------

def positions_to_symbols(beta: Sequence[int], moves: Sequence[Move]) -> List[Move]:
    flat = [x for mv in moves for x in mv]
    pos = (ctypes.c_int32 * max(1, len(flat)))(*flat)
    sym = (ctypes.c_int32 * max(1, len(flat)))()
    m = msgbuf()
    check(lib().td16_positions_to_symbols(_u8(beta), len(beta), pos, len(moves), sym, m, len(m)),
          m)
    return [(sym[3 * q], sym[3 * q + 1], sym[3 * q + 2]) for q in range(len(moves))]


def validate(beta: Sequence[int], omega: Sequence[int], symbol_moves: Sequence[Move],
             max_moves: int) -> None:
    """Raises Td16Error (IllegalArgumentException in the reference) when invalid."""
    flat = [x for mv in symbol_moves for x in mv]
    arr = (ctypes.c_int32 * max(1, len(flat)))(*flat)
    m = msgbuf()
    check(lib().td16_validate(_u8(beta), _u8(omega), len(beta), arr, len(symbol_moves),
                              int(max_moves), m, len(m)), m)


def codec_key(omega: Sequence[int]) -> bytes:
    buf = (ctypes.c_uint8 * 32)()
    ln = ctypes.c_int32(0)
    m = msgbuf()
    check(lib().td16_codec_key(_u8(omega), len(omega), buf, 32, ctypes.byref(ln), m, len(m)), m)
    return bytes(buf[:ln.value])


def codec_value(symbol_moves: Sequence[Move]) -> bytes:
    flat = [x for mv in symbol_moves for x in mv]
    arr = (ctypes.c_int32 * max(1, len(flat)))(*flat)
    buf = (ctypes.c_uint8 * 64)()
    ln = ctypes.c_int32(0)
    m = msgbuf()
    check(lib().td16_codec_value(arr, len(symbol_moves), buf, 64, ctypes.byref(ln), m, len(m)), m)
    return bytes(buf[:ln.value])


def codec_decode_key(data: bytes, n: int) -> List[int]:
    om = (ctypes.c_uint8 * 32)()
    m = msgbuf()
    check(lib().td16_codec_decode_key(_u8(list(data)), len(data), n, om, m, len(m)), m)
    return list(om[:n])


def codec_decode_value(data: bytes) -> List[Move]:
    sym = (ctypes.c_int32 * 96)()
    ln = ctypes.c_int32(0)
    m = msgbuf()
    check(lib().td16_codec_decode_value(_u8(list(data)), len(data), sym, 32, ctypes.byref(ln),
                                        m, len(m)), m)
    return [(sym[3 * q], sym[3 * q + 1], sym[3 * q + 2]) for q in range(ln.value)]


@dataclass
class Certificates:
    """Fixed-stride certificate records of a batch (td16_certify_batch)."""
    keys: np.ndarray       # uint8 [count, CERT_KEY_BYTES]; key q uses (5 n_q + 7) // 8 bytes
    values: np.ndarray     # uint8 [count, CERT_VALUE_BYTES]
    value_len: np.ndarray  # int32 [count]; 0 = nothing certified
    status: np.ndarray     # int32 [count]; 0 certified, 1 nothing to certify, <0 rejected

    def key(self, q: int, n: int) -> bytes:
        return bytes(self.keys[q, :(5 * n + 7) // 8])

    def value(self, q: int) -> bytes:
        return bytes(self.values[q, :self.value_len[q]])


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def certify_batch(reqs: Sequence[Request], results: Sequence[Result],
                  threads: int = 0) -> Certificates:
    """Positions -> symbols -> replay -> codec records for a batch of search
    results, on host threads (GPUSortingSearch.java:229-239,
    CertificateValidator.java:14-50, CertificateCodec.java:22-70)."""
    n = len(reqs)
    if len(results) != n:
        raise ValueError("reqs and results differ in length")
    rarr = reqs if isinstance(reqs, ctypes.Array) else (Request * max(1, n))(*reqs)
    oarr = results if isinstance(results, ctypes.Array) else (Result * max(1, n))(*results)
    keys = np.zeros((max(1, n), _lib_mod.CERT_KEY_BYTES), dtype=np.uint8)
    vals = np.zeros((max(1, n), _lib_mod.CERT_VALUE_BYTES), dtype=np.uint8)
    vlen = np.zeros(max(1, n), dtype=np.int32)
    st = np.zeros(max(1, n), dtype=np.int32)
    m = msgbuf()
    check(lib().td16_certify_batch(rarr, oarr, n, int(threads), _ptr(keys, ctypes.c_uint8),
                                   _ptr(vals, ctypes.c_uint8), _ptr(vlen, ctypes.c_int32),
                                   _ptr(st, ctypes.c_int32), m, len(m)), m)
    return Certificates(keys[:n], vals[:n], vlen[:n], st[:n])


def result_fields(results: ctypes.Array) -> dict:
    """status, found, depth, split (int32) and expanded (uint64) of a result
    array as numpy arrays (a byte view, no per-result loop)."""
    count = len(results)
    raw = np.frombuffer(memoryview(results).cast("B"), dtype=np.uint8).reshape(
        count, ctypes.sizeof(Result))
    out = {}
    for field, dt in (("status", np.int32), ("found", np.int32), ("depth", np.int32),
                      ("split", np.int32), ("expanded", np.uint64)):
        off = getattr(Result, field).offset
        w = np.dtype(dt).itemsize
        out[field] = raw[:, off:off + w].copy().view(dt).reshape(count)
    return out


def concat_requests(chunks: Sequence[ctypes.Array]) -> ctypes.Array:
    """One request array holding the given arrays back to back."""
    total = sum(len(c) for c in chunks)
    arr = (Request * max(1, total))()
    at = 0
    size = ctypes.sizeof(Request)
    for c in chunks:
        ctypes.memmove(ctypes.addressof(arr) + at * size, ctypes.addressof(c), len(c) * size)
        at += len(c)
    return arr if total else (Request * 0)()


def obligations(reqs: ctypes.Array):
    """(betas, omegas, ns, max_moves) of a request array as numpy arrays, the
    per-obligation inputs of check_batch (a byte view, no per-request loop)."""
    count = len(reqs)
    raw = np.frombuffer(memoryview(reqs).cast("B"), dtype=np.uint8).reshape(
        count, ctypes.sizeof(Request))

    def i32(field):
        off = getattr(Request, field).offset
        return raw[:, off:off + 4].copy().view(np.int32).reshape(count)

    b0, o0 = Request.beta.offset, Request.omega.offset
    mx = _lib_mod.MAX_N
    return (raw[:, b0:b0 + mx].copy(), raw[:, o0:o0 + mx].copy(), i32("n"), i32("max_moves"))


def check_batch(betas: np.ndarray, omegas: np.ndarray, ns: np.ndarray, max_moves,
                values: np.ndarray, value_len: np.ndarray, threads: int = 0) -> np.ndarray:
    """Decodes and replays a batch of stored certificates against their proof
    obligations (checker/.../CertificateStore.java:97-122).  betas, omegas:
    uint8 [count, MAX_N]; values: uint8 [count, CERT_VALUE_BYTES]; value_len 0
    means no stored value.  Returns int32 status per obligation (0 valid, 1 no
    value, -1..-6 replay rejection, -7 malformed value)."""
    count = len(ns)
    b = np.ascontiguousarray(betas, dtype=np.uint8).reshape(-1, _lib_mod.MAX_N)
    o = np.ascontiguousarray(omegas, dtype=np.uint8).reshape(-1, _lib_mod.MAX_N)
    v = np.ascontiguousarray(values, dtype=np.uint8).reshape(-1, _lib_mod.CERT_VALUE_BYTES)
    nn = np.ascontiguousarray(ns, dtype=np.int32)
    mm = np.ascontiguousarray(np.broadcast_to(np.asarray(max_moves, dtype=np.int32), (count,)))
    vl = np.ascontiguousarray(value_len, dtype=np.int32)
    if not (len(b) >= count and len(o) >= count and len(v) >= count and len(vl) >= count):
        raise ValueError("check_batch: arrays shorter than ns")
    st = np.zeros(max(1, count), dtype=np.int32)
    m = msgbuf()
    check(lib().td16_check_batch(_ptr(b, ctypes.c_uint8), _ptr(o, ctypes.c_uint8),
                                 _ptr(nn, ctypes.c_int32), _ptr(mm, ctypes.c_int32), count,
                                 _ptr(v, ctypes.c_uint8), _ptr(vl, ctypes.c_int32), int(threads),
                                 _ptr(st, ctypes.c_int32), m, len(m)), m)
    return st[:count]


# ---------------------------------------------------------------------------
# producer side: cycle-configuration representatives
# ---------------------------------------------------------------------------

def _bracelet_args(parts, oriented):
    m = len(parts)
    if oriented is None:
        oriented = [False] * m
    if len(oriented) != m:
        raise Td16Error(_lib_mod.INVALID_ARGUMENT,
                        "Partition and orientation arrays must have the same length")
    return (m, (ctypes.c_int32 * max(1, m))(*[int(p) for p in parts]),
            (ctypes.c_uint8 * max(1, m))(*[1 if o else 0 for o in oriented]))


def bracelet_count(parts: Sequence[int], oriented: Optional[Sequence[bool]] = None,
                   threads: int = 0, words: bool = False) -> int:
    """Number of realizable representatives (countRepresentatives,
    DirectTwistedBraceletGenerator.java:94-101), or of raw words."""
    m, P, O = _bracelet_args(parts, oriented)
    total = ctypes.c_int64(0)
    msg = msgbuf()
    mode = _lib_mod.BRACELET_WORDS if words else _lib_mod.BRACELET_PAIRS
    check(lib().td16_bracelets(P, O, m, mode, int(threads), None, 0, ctypes.byref(total),
                               msg, len(msg)), msg)
    return total.value


def bracelets(parts: Sequence[int], oriented: Optional[Sequence[bool]] = None,
              threads: int = 0, words: bool = False) -> np.ndarray:
    """Representatives of an annotated partition as a uint8 array [rows, n]:
    omega's one-line notation of each pair (omega, beta = (0 1 .. n-1)), or
    with words=True the raw orbit-minimal words (generateWords)."""
    m, P, O = _bracelet_args(parts, oriented)
    n = int(sum(int(p) for p in parts))
    total = ctypes.c_int64(0)
    mode = _lib_mod.BRACELET_WORDS if words else _lib_mod.BRACELET_PAIRS
    cap = 1 << 12
    while True:
        out = np.zeros((cap, max(1, n)), dtype=np.uint8)
        msg = msgbuf()
        st = lib().td16_bracelets(P, O, m, mode, int(threads), _ptr(out, ctypes.c_uint8), cap,
                                  ctypes.byref(total), msg, len(msg))
        if st == _lib_mod.OK:
            return out[:total.value]
        if total.value > cap:
            cap = total.value
            continue
        check(st, msg)


@dataclass
class BraceletClass:
    """The class of a pair (TwistedBraceletKey.of): annotated type, class
    word and representative pair (omega for beta = (0 1 .. n-1))."""
    parts: List[int]
    oriented: List[bool]
    word: Tuple[int, ...]
    representative: List[int]


def bracelet_canonical(beta: Sequence[int], omega: Sequence[int]) -> BraceletClass:
    n = len(beta)
    if len(omega) != n:
        raise Td16Error(_lib_mod.INVALID_ARGUMENT, "beta and omega differ in length")
    parts = (ctypes.c_int32 * 16)()
    ori = (ctypes.c_uint8 * 16)()
    m = ctypes.c_int32(0)
    word = (ctypes.c_uint8 * max(1, n))()
    rep = (ctypes.c_uint8 * max(1, n))()
    msg = msgbuf()
    check(lib().td16_bracelet_canonical(_u8(beta), _u8(omega), n, parts, ori, ctypes.byref(m),
                                        word, rep, msg, len(msg)), msg)
    k = m.value
    return BraceletClass(list(parts[:k]), [bool(x) for x in ori[:k]], tuple(word[:n]),
                         list(rep[:n]))


def bracelet_requests(omegas: np.ndarray, target: int, max_moves: int,
                      parts: Optional[Sequence[int]] = None) -> ctypes.Array:
    """Requests of the emitter's unoriented stage for representative pairs
    (UnorientedCycles.java:33-58 via WitnessSearch.java:23-34): beta = the
    identity cycle, omega, initial_even = omega's odd-length cycles (the odd
    parts when the rows' partition `parts` is given), min_rate =
    2 / (max_moves / (float) target), full sorting, REF, FIRST."""
    rows, n = omegas.shape
    if rows == 0:
        return (Request * 0)()
    arr = (Request * rows)()
    raw = np.frombuffer(memoryview(arr).cast("B"), dtype=np.uint8).reshape(
        rows, ctypes.sizeof(Request))
    b0, o0 = Request.beta.offset, Request.omega.offset
    raw[:, b0:b0 + n] = np.arange(n, dtype=np.uint8)
    raw[:, o0:o0 + n] = omegas
    if parts is not None:
        odd = sum(1 for p in parts if p % 2 == 1)
    else:
        # odd-length cycles of each omega, by walking all rows at once
        seen = np.zeros((rows, n), dtype=bool)
        odd = np.zeros(rows, dtype=np.int32)
        idx = np.arange(rows)
        for start in range(n):
            todo = ~seen[:, start]
            x = np.full(rows, start)
            length = np.zeros(rows, dtype=np.int32)
            while True:
                active = todo & ~seen[idx, x]
                if not active.any():
                    break
                seen[idx[active], x[active]] = True
                length[active] += 1
                x = np.where(active, omegas[idx, x].astype(np.int64), x)
            odd += (todo & (length % 2 == 1)).astype(np.int32)
    rate = np.float32(2) / (np.float32(max_moves) / np.float32(target))

    def put(field, values, dtype):
        off = getattr(Request, field).offset
        raw[:, off:off + 4] = np.ascontiguousarray(
            np.broadcast_to(np.asarray(values, dtype=dtype), (rows,))).view(np.uint8).reshape(rows, 4)

    put("n", n, np.int32)
    put("initial_even", odd, np.int32)
    put("min_rate", rate, np.float32)
    put("max_moves", max_moves, np.int32)
    put("full_sorting", 1, np.int32)
    return arr


# ---------------------------------------------------------------------------
# synthetic instances
# ---------------------------------------------------------------------------

def gen_uniform(seed: int, n: int, count: int) -> np.ndarray:
    out = np.zeros(count, dtype=np.uint64)
    st = lib().td16_gen_uniform(seed, n, count, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    if st != 0:
        raise Td16Error(st, "invalid generator arguments")
    return out


def gen_products(seed: int, nmin: int, nmax: int, kmin: int, kmax: int, count: int):
    ns = np.zeros(count, dtype=np.uint8)
    ks = np.zeros(count, dtype=np.uint8)
    perms = np.zeros(count, dtype=np.uint64)
    st = lib().td16_gen_products(seed, nmin, nmax, kmin, kmax, count,
                                 ns.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                 ks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                 perms.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    if st != 0:
        raise Td16Error(st, "invalid generator arguments")
    return ns, ks, perms


def gen_reverse_like(seed: int, n: int, count: int, rmax: int = 2) -> np.ndarray:
    """Nibble-packed permutations near the reverse permutation n-1 .. 0: the
    reverse followed by r ~ U{0..rmax} block transpositions (positions
    i < j < k uniform over 0..n, as in gen_products).  The reverse of 16 has
    the maximum distance 9 at that size; its neighbours are the hard cases of
    an emitter-shaped stream.  Python's Mersenne Twister seeded with `seed`."""
    import random
    rng = random.Random(seed)
    out = np.zeros(count, dtype=np.uint64)
    for c in range(count):
        p = list(range(n - 1, -1, -1))
        for _ in range(rng.randint(0, rmax)):
            i, j, k = sorted(rng.sample(range(n + 1), 3))
            p = p[:i] + p[j:k] + p[i:j] + p[k:]
        v = 0xFFFFFFFFFFFFFFFF
        for q in range(n):
            v = (v & ~(0xF << (4 * q))) | (p[q] << (4 * q))
        out[c] = v
    return out


def gen_emitter_stream(seed: int, count: int, reverse_every: int = 10):
    """An emitter-shaped stream of n = 16 instances, depth bound 9: every
    `reverse_every`-th instance is reverse-like (gen_reverse_like), the rest
    are products of 9 random block transpositions (gen_products).  Returns
    (ns, ks, perms) like gen_products."""
    hard = np.arange(count) % reverse_every == reverse_every - 1
    nh = int(hard.sum())
    ns = np.full(count, 16, dtype=np.uint8)
    ks = np.full(count, 9, dtype=np.uint8)
    perms = np.zeros(count, dtype=np.uint64)
    _, _, easy = gen_products(seed, 16, 16, 9, 9, count - nh)
    perms[~hard] = easy
    perms[hard] = gen_reverse_like(seed + 1, 16, nh)
    return ns, ks, perms


def gen_exponential(seed: int, count: int) -> np.ndarray:
    out = np.zeros(count, dtype=np.float64)
    lib().td16_gen_exponential(seed, count, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def version() -> str:
    return lib().td16_version().decode()


__all__ = [n for n in dir() if not n.startswith("_")] + ["FIRST", "COUNT_ALL", "MOVES_REF",
                                                        "MOVES_ALL"]
