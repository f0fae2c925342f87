"""The only route from a native result to a certificate (mirror of
emitter/src/main/java/br/unb/cic/tdp/base/WitnessSearch.java:23-48 and
SortingSearch.java:16-20): search, replay-validate, then hand the encoded
certificate to a sink.

The reference's sink is a RocksDB writer (AsyncCertificateWriter.java); this
package defines the sink interface and an in-memory sink, keyed and encoded
with the reference's certificate codec (CertificateCodec.java:22-118).
"""
from __future__ import annotations

import threading
from typing import Dict, List, Optional, Sequence, Tuple

from . import native
from .gpu_sorting_search import GPUSortingSearch

Move = Tuple[int, int, int]


class MemoryCertificateSink:
    """key = codec_key(omega), value = codec_value(symbol moves)."""

    def __init__(self):
        self._lock = threading.Lock()
        self.entries: Dict[bytes, bytes] = {}

    def enqueue(self, omega: Sequence[int], moves: Sequence[Move]) -> None:
        key = native.codec_key(omega)
        value = native.codec_value(moves)
        with self._lock:
            self.entries[key] = value

    def find(self, omega: Sequence[int]) -> Optional[List[Move]]:
        v = self.entries.get(native.codec_key(omega))
        return None if v is None else native.codec_decode_value(v)


class WitnessSearch:
    _writer = None

    @classmethod
    def install_writer(cls, writer) -> None:
        if cls._writer is not None:
            raise RuntimeError("A proof-scoped certificate writer is already installed")
        cls._writer = writer

    @classmethod
    def uninstall_writer(cls) -> None:
        cls._writer = None

    @classmethod
    def search(cls, beta: Sequence[int], omega: Sequence[int], minimum_rate: float,
               maximum_moves: int) -> List[Move]:
        # initialEvenCycles = omega.getNumberOfEvenCycles(): odd-length cycles, fixed points included
        initial_even = _odd_cycles(omega)
        moves = GPUSortingSearch.get_instance().search(beta, omega, initial_even, minimum_rate,
                                                       maximum_moves, True)
        if not moves:
            return moves
        # the rate configures the search; it is not a certificate obligation
        native.validate(beta, omega, moves, maximum_moves)
        if cls._writer is None:
            raise RuntimeError("Native witness succeeded without a proof-scoped certificate writer")
        cls._writer.enqueue(omega, moves)
        return moves


def _odd_cycles(omega: Sequence[int]) -> int:
    seen = [False] * len(omega)
    odd = 0
    for s in range(len(omega)):
        if seen[s]:
            continue
        ln = 0
        x = s
        while not seen[x]:
            seen[x] = True
            x = omega[x]
            ln += 1
        odd += ln & 1
    return odd
