"""ctypes view of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only to check or time the reference restatement in
``oracle/td16_oracle.c``.  The product package ``td16_amd`` never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libtd16_oracle.so")

MOVES_REF, MOVES_ALL = 0, 1
FIRST, COUNT_ALL = 0, 1


class Result(ctypes.Structure):
    _fields_ = [
        ("found", ctypes.c_int32),
        ("depth", ctypes.c_int32),
        ("moves", (ctypes.c_int32 * 3) * 9),
        ("expanded", ctypes.c_uint64),
        ("solutions", ctypes.c_uint64),
    ]

    def witness(self):
        return [tuple(self.moves[d]) for d in range(self.depth)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        i32p = ctypes.POINTER(ctypes.c_int32)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.td16o_count_odd_cycles.argtypes = [u8p, ctypes.c_int]
        L.td16o_delta.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.td16o_rate_tables.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                        ctypes.c_int, ctypes.c_int, u32p, u32p]
        L.td16o_rate_tables.restype = None
        L.td16o_search.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(Result)]
        L.td16o_search_order.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.POINTER(Result)]
        L.td16o_iddfs_order.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.POINTER(Result)]
        L.td16o_search_tables.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, u32p, u32p,
                                          ctypes.c_int, ctypes.c_int, ctypes.POINTER(Result)]
        L.td16o_iddfs.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(Result)]
        L.td16o_count_naive.argtypes = [u8p, u8p, ctypes.c_int, ctypes.c_int, u32p, u32p,
                                        ctypes.c_int, u64p, u64p]
        L.td16o_positions_to_symbols.argtypes = [u8p, ctypes.c_int, i32p, ctypes.c_int, i32p]
        L.td16o_validate.argtypes = [u8p, u8p, ctypes.c_int, i32p, ctypes.c_int, ctypes.c_int]
        L.td16o_codec_key.argtypes = [u8p, ctypes.c_int, u8p, ctypes.c_int]
        L.td16o_codec_value.argtypes = [i32p, ctypes.c_int, u8p, ctypes.c_int]
        L.td16o_codec_decode_value.argtypes = [u8p, ctypes.c_int, i32p, ctypes.c_int]
        L.td16o_codec_decode_key.argtypes = [u8p, ctypes.c_int, ctypes.c_int, u8p]
        L.td16o_splitmix64.argtypes = [u64p]
        L.td16o_splitmix64.restype = ctypes.c_uint64
        L.td16o_uniform.argtypes = [u64p, ctypes.c_uint64]
        L.td16o_uniform.restype = ctypes.c_uint64
        L.td16o_gen_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, u64p]
        L.td16o_gen_uniform.restype = None
        L.td16o_gen_products.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         u8p, u8p, u64p]
        L.td16o_gen_products.restype = None
        L.td16o_linear_to_pair.argtypes = [ctypes.c_uint64, ctypes.c_int, u8p, u8p]
        L.td16o_bfs.argtypes = [ctypes.c_int, u8p, u64p]
        L.td16o_perm_rank.argtypes = [u8p, ctypes.c_int]
        L.td16o_perm_rank.restype = ctypes.c_uint64
        L.td16o_perm_unrank.argtypes = [ctypes.c_uint64, ctypes.c_int, u8p]
        L.td16o_perm_unrank.restype = None
        L.td16o_search_batch.argtypes = [ctypes.c_int, u8p, u64p, u8p, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(Result)]
        L.td16o_search_batch_ex.argtypes = [ctypes.c_int, u8p, u64p, u8p, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_uint64, ctypes.c_int,
                                            ctypes.POINTER(Result)]
        _lib = L
    return _lib


def _u8(seq):
    a = (ctypes.c_uint8 * max(1, len(seq)))(*seq)
    return a


def count_odd_cycles(omega):
    return lib().td16o_count_odd_cycles(_u8(omega), len(omega))


def delta(omega, a, b, c):
    d = lib().td16o_delta(_u8(omega), a, b, c)
    return None if d == -1000 else d


def rate_tables(n, initial_even, min_rate, max_depth, full_sorting):
    sat = (ctypes.c_uint32 * 10)()
    reach = (ctypes.c_uint32 * 10)()
    lib().td16o_rate_tables(n, initial_even, min_rate, max_depth, int(full_sorting), sat, reach)
    return list(sat), list(reach)


def search(beta, omega, initial_even, min_rate, max_depth, full_sorting=True,
           moveset=MOVES_REF, mode=FIRST, greedy=False):
    """greedy=False: the reference's lexicographic DFS; greedy=True: children
    gaining odd cycles first (the kernel's default FIRST order)."""
    r = Result()
    rc = lib().td16o_search_order(_u8(beta), _u8(omega), len(beta), initial_even, min_rate,
                                  max_depth, int(full_sorting), moveset, mode, int(greedy),
                                  ctypes.byref(r))
    if rc != 0:
        raise ValueError("invalid search input")
    return r


def search_tables(beta, omega, max_depth, sat, reach, moveset=MOVES_REF, mode=FIRST):
    r = Result()
    rc = lib().td16o_search_tables(_u8(beta), _u8(omega), len(beta), max_depth,
                                   (ctypes.c_uint32 * 10)(*sat), (ctypes.c_uint32 * 10)(*reach),
                                   moveset, mode, ctypes.byref(r))
    if rc != 0:
        raise ValueError("invalid search input")
    return r


def iddfs(beta, omega, moveset=MOVES_REF, max_limit=9, greedy=False):
    r = Result()
    d = lib().td16o_iddfs_order(_u8(beta), _u8(omega), len(beta), moveset, max_limit,
                                int(greedy), ctypes.byref(r))
    return d, r


def count_naive(beta, omega, max_depth, sat, reach, moveset):
    e = ctypes.c_uint64()
    s = ctypes.c_uint64()
    rc = lib().td16o_count_naive(_u8(beta), _u8(omega), len(beta), max_depth,
                                 (ctypes.c_uint32 * 10)(*sat), (ctypes.c_uint32 * 10)(*reach),
                                 moveset, ctypes.byref(e), ctypes.byref(s))
    if rc != 0:
        raise ValueError("invalid input")
    return e.value, s.value


def positions_to_symbols(beta, moves):
    flat = [x for m in moves for x in m]
    pos = (ctypes.c_int32 * max(1, len(flat)))(*flat)
    sym = (ctypes.c_int32 * max(1, len(flat)))()
    rc = lib().td16o_positions_to_symbols(_u8(beta), len(beta), pos, len(moves), sym)
    if rc != 0:
        raise ValueError("invalid positions")
    return [tuple(sym[3 * m:3 * m + 3]) for m in range(len(moves))]


def validate(beta, omega, sym_moves, max_moves):
    flat = [x for m in sym_moves for x in m]
    arr = (ctypes.c_int32 * max(1, len(flat)))(*flat)
    return lib().td16o_validate(_u8(beta), _u8(omega), len(beta), arr, len(sym_moves),
                                max_moves)


def codec_key(omega):
    buf = (ctypes.c_uint8 * 32)()
    nb = lib().td16o_codec_key(_u8(omega), len(omega), buf, 32)
    if nb < 0:
        raise ValueError("bad key input")
    return bytes(buf[:nb])


def codec_value(sym_moves):
    flat = [x for m in sym_moves for x in m]
    arr = (ctypes.c_int32 * max(1, len(flat)))(*flat)
    buf = (ctypes.c_uint8 * 64)()
    nb = lib().td16o_codec_value(arr, len(sym_moves), buf, 64)
    if nb < 0:
        raise ValueError("bad value input")
    return bytes(buf[:nb])


def codec_decode_value(data: bytes):
    buf = _u8(list(data))
    sym = (ctypes.c_int32 * 96)()
    m = lib().td16o_codec_decode_value(buf, len(data), sym, 32)
    if m < 0:
        raise ValueError("malformed certificate value")
    return [tuple(sym[3 * q:3 * q + 3]) for q in range(m)]


def codec_decode_key(data: bytes, n):
    buf = _u8(list(data))
    om = (ctypes.c_uint8 * 32)()
    if lib().td16o_codec_decode_key(buf, len(data), n, om) != 0:
        raise ValueError("malformed certificate key")
    return list(om[:n])


def splitmix64_stream(seed, count):
    st = ctypes.c_uint64(seed)
    return [lib().td16o_splitmix64(ctypes.byref(st)) for _ in range(count)]


def gen_uniform(seed, n, count):
    out = np.zeros(count, dtype=np.uint64)
    lib().td16o_gen_uniform(seed, n, count,
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return out


def gen_products(seed, nmin, nmax, kmin, kmax, count):
    ns = np.zeros(count, dtype=np.uint8)
    ks = np.zeros(count, dtype=np.uint8)
    perms = np.zeros(count, dtype=np.uint64)
    lib().td16o_gen_products(seed, nmin, nmax, kmin, kmax, count,
                             ns.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                             ks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                             perms.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return ns, ks, perms


def linear_to_pair(perm, n):
    beta = (ctypes.c_uint8 * 17)()
    omega = (ctypes.c_uint8 * 17)()
    N = lib().td16o_linear_to_pair(int(perm), n, beta, omega)
    return list(beta[:N]), list(omega[:N])


def unpack(perm, n):
    return [(int(perm) >> (4 * q)) & 0xF for q in range(n)]


def pack(p):
    v = (1 << 64) - 1
    for q, x in enumerate(p):
        v &= ~(0xF << (4 * q))
        v |= x << (4 * q)
    return v


def bfs(n, with_paths=False):
    import math
    total = math.factorial(n)
    dist = np.zeros(total, dtype=np.uint8)
    paths = np.zeros(total, dtype=np.uint64) if with_paths else None
    diam = lib().td16o_bfs(n, dist.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                           paths.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
                           if with_paths else None)
    return diam, dist, paths


def perm_unrank(r, n):
    p = (ctypes.c_uint8 * 16)()
    lib().td16o_perm_unrank(r, n, p)
    return list(p[:n])


def perm_rank(p):
    return lib().td16o_perm_rank(_u8(p), len(p))


def search_batch(ns, perms, ks, moveset=MOVES_REF, mode=FIRST, threads=1, greedy=False):
    count = len(ns)
    out = (Result * count)()
    ns = np.ascontiguousarray(ns, dtype=np.uint8)
    perms = np.ascontiguousarray(perms, dtype=np.uint64)
    ks = np.ascontiguousarray(ks, dtype=np.uint8)
    lib().td16o_search_batch(count, ns.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                             perms.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                             ks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                             moveset, mode, int(greedy), threads, out)
    return out


def search_batch_ex(ns, perms, ks, moveset=MOVES_REF, mode=FIRST, threads=1, greedy=False,
                    iddfs=False, cap=0):
    """search_batch with iterative deepening (exact distance within k) and a
    node cap (found = -2 once a search passes `cap` expanded nodes)."""
    count = len(ns)
    out = (Result * max(1, count))()
    ns = np.ascontiguousarray(ns, dtype=np.uint8)
    perms = np.ascontiguousarray(perms, dtype=np.uint64)
    ks = np.ascontiguousarray(ks, dtype=np.uint8)
    lib().td16o_search_batch_ex(count, ns.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                perms.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                ks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                moveset, mode, int(greedy), int(iddfs), int(cap), threads, out)
    return out
