"""CPU tests of the producer side: td16_bracelets (one representative per
realizable cycle-configuration class) against brute-force orbits
(oracle/bracelets.py), the reference generator's test fixtures and the class
counts of the emitter's unoriented stage (tests/golden/bracelets.json)."""
import json
import os

import numpy as np
import pytest

from oracle import bracelets as B
from td16_amd import Td16Error
from td16_amd import _lib

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def gold():
    with open(os.path.join(GOLDEN, "bracelets.json")) as f:
        return json.load(f)


def cycle_string(omega):
    return "".join("(" + " ".join(map(str, c)) + ")" for c in B.cycles([int(x) for x in omega]))


def generated_classes(native, parts, oriented):
    """Class words of the generator's pairs, with checks on each pair."""
    rows = native.bracelets(parts, oriented)
    want = tuple(sorted(zip(parts, oriented)))
    out = set()
    for row in rows:
        omega = [int(x) for x in row]
        assert sorted(omega) == list(range(len(omega)))
        assert all(omega[x] != x for x in range(len(omega)))
        assert B.target_is_one_cycle(omega)
        key, word, A = B.annotated_type(omega)
        assert key == want
        out.add(B.class_word(word, A))
    assert len(out) == len(rows), "two pairs of one class"
    return out


@pytest.mark.parametrize("parts,oriented", [
    ([5, 3], [False, True]), ([5, 2, 2], [False] * 3), ([4, 2, 2], [False] * 3),
    ([2, 5, 2], [False] * 3), ([4, 3], [True, True]), ([3, 2, 2], [True, False, False]),
    ([2, 4, 2], [False, True, False]), ([3, 3], [False, False]), ([3, 3], [True, False]),
    ([6], [True]), ([5], [True])])
def test_raw_words_equal_brute_force_orbits(native, parts, oriented):
    # DirectTwistedBraceletGeneratorTest.java:216-245: the raw orbit-minimal words
    got = [tuple(int(x) for x in r) for r in native.bracelets(parts, oriented, words=True)]
    assert len(got) == len(set(got))
    assert set(got) == B.raw_words(list(zip(parts, oriented)))
    assert got == sorted(got)   # increasing word order


def test_raw_word_fixtures(native):
    six = [list(r) for r in native.bracelets([6], [True], words=True)]
    assert len(six) == 19 and [0, 5, 4, 3, 2, 1] not in six        # :184-195
    three = [list(r) for r in native.bracelets([3, 3], [False, False], words=True)]
    assert [0, 0, 1, 1, 0, 1] not in three                           # :196-204
    long = native.bracelets([24], [False], words=True)               # :206-214
    assert long.shape == (1, 24) and not long.any()


@pytest.mark.parametrize("parts,oriented", [
    ([3, 3], [False, False]), ([3, 3], [True, False]), ([3, 3], [True, True]),
    ([5], [True]), ([6], [True]), ([5, 3], [True, False])])
def test_small_families_equal_brute_force(native, parts, oriented):
    # DirectTwistedBraceletGeneratorTest.java:32-46 against every n-cycle target
    n = sum(parts)
    expected = B.classes_by_pair(n).get(tuple(sorted(zip(parts, oriented))), set())
    assert generated_classes(native, parts, oriented) == expected


@pytest.mark.parametrize("n", [8, 9])
def test_every_annotated_type(native, gold, n):
    # DirectTwistedBraceletGeneratorTest.java:246-291: every fixed-point-free
    # annotated type of S_n (17 at n = 8), realizable or not
    types = B.annotated_types(n)
    if str(n) in gold["annotated_type_counts"]:
        assert len(types) == gold["annotated_type_counts"][str(n)]
    assert len(B.annotated_types(11)) == gold["annotated_type_counts"]["11"]
    expected = B.classes_by_pair(n)
    assert set(expected) <= types
    for t in types:
        parts = [s for s, _ in t]
        oriented = [o for _, o in t]
        assert generated_classes(native, parts, oriented) == expected.get(t, set()), t


def test_reference_counts_and_persistent_representative(native, gold):
    for case in gold["test_counts"]:
        assert native.bracelet_count(case["parts"], case["oriented"]) == case["classes"], case
    rep = gold["persistent_representative"]
    omegas = {cycle_string(r) for r in native.bracelets(rep["parts"])}
    assert len(omegas) == 4800
    assert rep["present"] in omegas and rep["absent"] not in omegas


def test_unoriented_stage_counts_of_the_emitter_log(native, gold):
    # emitter.log:28-68: every unoriented partition of 17 in the stage, 3,858,884
    # classes in all
    total = 0
    for case in gold["unoriented_stage"]:
        got = native.bracelet_count(case["parts"], None, threads=0)
        assert got == case["classes"], case
        total += got
    assert total == 3858884


def test_rows_do_not_depend_on_threads(native):
    for parts in ([7, 4, 2, 2, 2], [6, 4, 3, 2, 2]):
        a = native.bracelets(parts, threads=1)
        b = native.bracelets(parts, threads=5)
        assert np.array_equal(a, b)
    a = native.bracelets([5, 3, 3], [True, False, False], threads=1, words=True)
    assert np.array_equal(a, native.bracelets([5, 3, 3], [True, False, False], threads=3,
                                               words=True))


def test_stage_requests(native):
    # UnorientedCycles.java:33-58 -> WitnessSearch.java:23-34
    rows = native.bracelets([7, 4, 2, 2, 2])
    reqs = native.bracelet_requests(rows, target=8, max_moves=9)
    assert len(reqs) == 4800
    for q in (0, 1234, 4799):
        r = reqs[q]
        assert r.n == 17 and list(r.beta) == list(range(17))
        assert list(r.omega) == [int(x) for x in rows[q]]
        assert r.initial_even == 1                 # one odd-length cycle: the 7
        assert r.min_rate == float(np.float32(2) / (np.float32(9) / np.float32(8)))
        assert r.max_moves == 9 and r.full_sorting == 1 and r.mode == 0 and r.moveset == 0
    rows = native.bracelets([5, 3, 3, 3, 3])
    reqs = native.bracelet_requests(rows[:10], target=6, max_moves=9)
    assert all(reqs[q].initial_even == 5 for q in range(10))


def test_invalid_partitions(native):
    # DirectTwistedBraceletGeneratorTest.java:48-65
    for parts, oriented in (([], []), ([3], [True, False]), ([0], [False]), ([1], [False]),
                            ([2], [True])):
        with pytest.raises(Td16Error) as e:
            native.bracelet_count(parts, oriented)
        assert e.value.status == _lib.INVALID_ARGUMENT
    assert native.bracelet_count([2], [False]) == 0
    with pytest.raises(Td16Error):
        native.bracelet_count([20, 13])          # more than 32 beads


def test_canonical_form_is_the_generators_row(native):
    # canonicalRepresentative (TwistedBraceletKey.java:96-98) of each
    # generated pair is that pair, also when beta starts elsewhere on its
    # circle (a rotated word) and when the symbols are renamed along beta
    rng = np.random.default_rng(7)
    for parts, oriented in (([7, 4, 2, 2, 2], None), ([5, 3, 3], [True, False, False]),
                            ([4, 3], [True, True]), ([6, 6, 5], None)):
        rows = native.bracelets(parts, oriented)
        words = native.bracelets(parts, oriented, words=True)
        want = sorted(zip(parts, oriented or [False] * len(parts)), key=lambda a: (-a[0], a[1]))
        word_set = {tuple(int(x) for x in w) for w in words}
        for q in rng.choice(len(rows), size=min(200, len(rows)), replace=False):
            omega = [int(x) for x in rows[q]]
            n = len(omega)
            cls = native.bracelet_canonical(list(range(n)), omega)
            assert list(zip(cls.parts, cls.oriented)) == want
            assert cls.representative == omega
            assert cls.word in word_set
            s = int(rng.integers(1, n))
            beta = [(p + s) % n for p in range(n)]          # same cyclic sequence
            omega2 = [0] * n
            for x in range(n):                               # omega relabelled along beta
                omega2[beta[x]] = beta[omega[x]]
            assert native.bracelet_canonical(beta, omega2).representative == omega
            assert native.bracelet_canonical(beta, omega).representative == omega


def test_canonical_form_matches_brute_force_classes(native):
    rng = np.random.default_rng(11)
    done = 0
    while done < 400:
        n = int(rng.integers(4, 10))
        omega = [int(x) for x in rng.permutation(n)]
        if any(omega[x] == x for x in range(n)):
            continue
        beta = [int(x) for x in rng.permutation(n)]
        # the same pair read along beta: positions instead of symbols
        pos = {b: p for p, b in enumerate(beta)}
        om = [pos[omega[beta[p]]] for p in range(n)]
        key, word, A = B.annotated_type(om)
        cls = native.bracelet_canonical(beta, omega)
        assert tuple(sorted(zip(cls.parts, cls.oriented))) == key
        assert cls.word == B.class_word(word, A)
        assert cls.representative == B.decode(cls.word, A)
        done += 1
    with pytest.raises(Td16Error):
        native.bracelet_canonical([0, 1, 2], [0, 2, 1])     # fixed point
    with pytest.raises(Td16Error):
        native.bracelet_canonical([0, 1, 1], [1, 2, 0])     # not a permutation


def test_small_output_buffer_reports_the_count(native):
    import ctypes
    parts = (ctypes.c_int32 * 5)(6, 4, 3, 2, 2)
    ori = (ctypes.c_uint8 * 5)()
    for threads in (1, 4):
        out = np.zeros((10, 17), np.uint8)
        total = ctypes.c_int64(0)
        msg = _lib.msgbuf()
        st = native.lib().td16_bracelets(parts, ori, 5, _lib.BRACELET_PAIRS, threads,
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                         10, ctypes.byref(total), msg, len(msg))
        assert st == _lib.INVALID_ARGUMENT and total.value == 119000
        assert b"119000" in msg.value
        rows = native.bracelets([6, 4, 3, 2, 2])          # the wrapper retries with room
        assert {bytes(r) for r in out} <= {bytes(r) for r in rows}
