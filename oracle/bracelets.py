"""Brute-force cycle-configuration classes -- TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, to check the product's generator
(``td16_bracelets``) against definitions that share no state with it:

- ``raw_words``: every fixed-content word of an annotated partition, the
  unoriented-pattern words of oriented cycles dropped, each replaced by its
  least standardized traversal over all rotations and reflections
  (emitter/src/test/java/br/unb/cic/tdp/DirectTwistedBraceletGeneratorTest.java:
  378-465, the reference's own brute force);
- ``classes_by_pair``: every pair (omega, beta = (0 1 .. n-1)) whose target
  omega o beta is one n-cycle, enumerated from the n-cycles themselves,
  grouped by annotated type and reduced to its class word
  (DirectTwistedBraceletGeneratorTest.java:256-291, 498-555, 573-598).

Pure Python; sized for n <= 9.
"""
from __future__ import annotations

import itertools
from typing import Dict, List, Sequence, Set, Tuple

Word = Tuple[int, ...]
Annotation = Tuple[int, bool]  # (size, oriented)


class Alphabet:
    """Symbols of an annotated partition: cycles longest first, unoriented
    before oriented of one size (DirectTwistedBraceletGenerator.java:257-270);
    one symbol per unoriented cycle, one per rank of an oriented cycle."""

    def __init__(self, annotations: Sequence[Annotation]):
        self.types: List[Annotation] = sorted(annotations, key=lambda a: (-a[0], a[1]))
        self.first: List[int] = []
        self.sym_cycle: List[int] = []
        self.sym_rank: List[int] = []
        self.mult: List[int] = []
        for c, (size, ori) in enumerate(self.types):
            self.first.append(len(self.sym_cycle))
            for r in range(size if ori else 1):
                self.sym_cycle.append(c)
                self.sym_rank.append(r)
                self.mult.append(1 if ori else size)
        self.n = sum(s for s, _ in self.types)


def standardized_traversal(word: Word, start: int, direction: int, A: Alphabet) -> Word:
    """DirectTwistedBraceletGeneratorTest.java:435-464."""
    n = len(word)
    target_by_cycle: Dict[int, int] = {}
    used: Set[int] = set()
    origin: Dict[int, int] = {}
    out = []
    for k in range(n):
        bead = word[(start + direction * k) % n]
        c = A.sym_cycle[bead]
        size, ori = A.types[c]
        if c not in target_by_cycle:
            t = next(t for t in range(len(A.types)) if t not in used and A.types[t] == A.types[c])
            target_by_cycle[c] = t
            used.add(t)
            origin[c] = A.sym_rank[bead]
        rank = (direction * (A.sym_rank[bead] - origin[c])) % size if ori else 0
        out.append(A.first[target_by_cycle[c]] + rank)
    return tuple(out)


def class_word(word: Word, A: Alphabet) -> Word:
    n = len(word)
    return min(standardized_traversal(word, s, d, A) for s in range(n) for d in (1, -1))


def unoriented_pattern(word: Word, A: Alphabet) -> bool:
    """An oriented cycle whose ranks read r, r-1, r-2, .. along the word
    (DirectTwistedBraceletGeneratorTest.java:399-418)."""
    for c, (size, ori) in enumerate(A.types):
        if not ori:
            continue
        prev = None
        descending = True
        for bead in word:
            if A.sym_cycle[bead] != c:
                continue
            if prev is not None and (A.sym_rank[bead] + 1) % size != prev:
                descending = False
            prev = A.sym_rank[bead]
        if descending:
            return True
    return False


def raw_words(annotations: Sequence[Annotation]) -> Set[Word]:
    A = Alphabet(annotations)
    pool = [s for s in range(len(A.mult)) for _ in range(A.mult[s])]
    out: Set[Word] = set()
    for word in set(itertools.permutations(pool)):
        if unoriented_pattern(word, A):
            continue
        out.add(class_word(word, A))
    return out


def decode(word: Word, A: Alphabet) -> List[int]:
    """omega of a word: an unoriented cycle met at p1 < .. < pk maps p(i) to
    p(i-1) and p1 to pk; an oriented one maps rank r's position to rank
    r+1's (DirectTwistedBraceletGenerator.java:1475-1505)."""
    n = len(word)
    omega = [0] * n
    for c, (size, ori) in enumerate(A.types):
        if ori:
            pos = {A.sym_rank[b]: p for p, b in enumerate(word) if A.sym_cycle[b] == c}
            for r in range(size):
                omega[pos[r]] = pos[(r + 1) % size]
        else:
            ps = [p for p, b in enumerate(word) if A.sym_cycle[b] == c]
            for i, p in enumerate(ps):
                omega[p] = ps[i - 1]
    return omega


def cycles(perm: Sequence[int]) -> List[List[int]]:
    seen = [False] * len(perm)
    out = []
    for s in range(len(perm)):
        if seen[s]:
            continue
        cyc, x = [], s
        while not seen[x]:
            seen[x] = True
            cyc.append(x)
            x = perm[x]
        out.append(cyc)
    return out


def target_is_one_cycle(omega: Sequence[int]) -> bool:
    n = len(omega)
    rho = [omega[(x + 1) % n] for x in range(n)]
    return len(cycles(rho)) == 1


def annotated_type(omega: Sequence[int]) -> Tuple[Tuple[Annotation, ...], Word, Alphabet]:
    """Annotated type of (omega, beta = identity) and the word that lists
    each position's cycle (and rank): a cycle is unoriented when omega maps
    each of its positions to the cycle's previous position along beta."""
    n = len(omega)
    cyc = cycles(omega)
    ann = []
    for c in cyc:
        ps = sorted(c)
        unori = all(omega[p] == ps[i - 1] for i, p in enumerate(ps))
        ann.append((len(c), not unori))
    A = Alphabet(ann)
    # assign each cycle to a slot of its type, in any order
    free = {}
    for slot, a in enumerate(A.types):
        free.setdefault(a, []).append(slot)
    word = [0] * n
    for c, a in zip(cyc, ann):
        slot = free[a].pop(0)
        if a[1]:
            x = c[0]
            for r in range(len(c)):
                word[x] = A.first[slot] + r
                x = omega[x]
        else:
            for p in c:
                word[p] = A.first[slot]
    return tuple(sorted(A.types)), tuple(word), A


def classes_by_pair(n: int) -> Dict[Tuple[Annotation, ...], Set[Word]]:
    """Every fixed-point-free omega with omega o beta one n-cycle, from the
    n-cycles rho (omega(x) = rho(x - 1)), grouped by annotated type; each
    reduced to its class word."""
    out: Dict[Tuple[Annotation, ...], Set[Word]] = {}
    for tail in itertools.permutations(range(1, n)):
        order = (0,) + tail
        rho = [0] * n
        for i in range(n):
            rho[order[i]] = order[(i + 1) % n]
        omega = [rho[(x - 1) % n] for x in range(n)]
        if any(omega[x] == x for x in range(n)):
            continue
        key, word, A = annotated_type(omega)
        out.setdefault(key, set()).add(class_word(word, A))
    return out


def annotated_types(n: int) -> Set[Tuple[Annotation, ...]]:
    """Every fixed-point-free annotated type on n symbols: partitions into
    parts >= 2, any number of each size >= 3 oriented
    (DirectTwistedBraceletGeneratorTest.java:293-343)."""
    out = set()

    def parts(rem, mx, acc):
        if rem == 0:
            groups = [(s, sum(1 for x in acc if x == s)) for s in sorted(set(acc))]
            choices = [range(0, k + 1) if s >= 3 else range(0, 1) for s, k in groups]
            for pick in itertools.product(*choices):
                ann = []
                for (s, k), o in zip(groups, pick):
                    ann += [(s, False)] * (k - o) + [(s, True)] * o
                out.add(tuple(sorted(ann)))
            return
        for p in range(min(rem, mx), 1, -1):
            parts(rem - p, p, acc + [p])

    parts(n, n, [])
    return out
