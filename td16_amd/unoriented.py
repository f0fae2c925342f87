"""The emitter's unoriented stage without a JVM: every realizable class of
each all-unoriented partition of 17 is searched for a sorting sequence of at
most 9 moves (emitter/src/main/java/br/unb/cic/tdp/WitnessEmitter.java:16-27,
60-65; UnorientedCycles.java:33-72, 93-98).

The reference walks the partitions one by one and sends each representative
through the bounded slot dispatcher; here a stage (or any subset of it) is
generated on host threads (td16_bracelets) and searched in one batched launch.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import native

MAX_MOVES = 9

# WitnessEmitter.java:17-28: partitions of 17 and the odd-cycle increase each
# must reach within 9 moves
P2_RATE_EIGHT = [
    [8, 7, 2], [7, 4, 2, 2, 2], [7, 6, 4], [6, 5, 2, 2, 2],
    [6, 4, 3, 2, 2], [6, 6, 5], [5, 4, 4, 2, 2],
    [5, 2, 2, 2, 2, 2, 2], [4, 4, 4, 3, 2], [4, 3, 2, 2, 2, 2, 2],
]
P2_RATE_SEVEN = [
    [8, 6, 3], [8, 5, 4], [7, 7, 3], [7, 3, 3, 2, 2], [7, 5, 5],
    [6, 3, 3, 3, 2], [5, 5, 3, 2, 2], [5, 4, 3, 3, 2],
    [4, 4, 3, 3, 3], [3, 3, 3, 2, 2, 2, 2],
]
P2_RATE_SIX = [[5, 3, 3, 3, 3]]

STAGE: List[Tuple[List[int], int]] = (
    [(p, 8) for p in P2_RATE_EIGHT] + [(p, 7) for p in P2_RATE_SEVEN]
    + [(p, 6) for p in P2_RATE_SIX])


def has_forced_target_fixed_point(parts: Sequence[int]) -> bool:
    """A cycle longer than all others together forces two adjacent beads of
    one unoriented cycle, hence a target fixed point: no class is realizable
    (UnorientedCycles.java:93-98)."""
    n = sum(parts)
    mx = max(parts) if parts else 0
    return n > 1 and mx > n - mx


@dataclass
class StageBatch:
    requests: object            # ctypes array of td16_request
    partition: np.ndarray       # index into `cases` of each request
    cases: List[Tuple[List[int], int]]
    classes: List[int]          # realizable classes per case


def stage_requests(cases: Optional[Sequence[Tuple[Sequence[int], int]]] = None,
                   threads: int = 0) -> StageBatch:
    """Requests for every realizable class of each (partition, target) case
    (default: the whole stage), in case order and, within a case, in the
    generator's word order."""
    cases = [(list(p), int(t)) for p, t in (STAGE if cases is None else cases)]
    chunks, index, classes = [], [], []
    for i, (parts, target) in enumerate(cases):
        if has_forced_target_fixed_point(parts):
            rows = np.zeros((0, sum(parts)), dtype=np.uint8)
        else:
            rows = native.bracelets(parts, None, threads)
        chunks.append(native.bracelet_requests(rows, target, MAX_MOVES, parts))
        index.append(np.full(len(rows), i, dtype=np.int32))
        classes.append(len(rows))
    reqs = native.concat_requests(chunks)
    part = np.concatenate(index) if index else np.zeros(0, dtype=np.int32)
    return StageBatch(reqs, part, cases, classes)
