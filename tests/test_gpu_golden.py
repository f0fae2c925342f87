"""GPU results against the committed oracle fixtures (tests/golden/), the
reference's known answers, and the edge cases its native layer handles."""
import json
import os
import threading

import numpy as np
import pytest

from td16_amd import GPUSortingSearch, MemoryCertificateSink, Td16Error, WitnessSearch
from td16_amd import _lib

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def exp():
    with open(os.path.join(GOLDEN, "expected.json")) as f:
        return json.load(f)


def pair(req):
    return list(req.beta[:req.n]), list(req.omega[:req.n])


def assert_matches(O, native, req, res, ref, exact=False):
    assert res.status == 0
    assert res.found == ref["found"]
    if res.found and res.depth:
        beta, omega = pair(req)
        sym = native.positions_to_symbols(beta, res.witness())
        assert O.validate(beta, omega, sym, req.max_moves) == 0
    if exact or not ref["found"]:
        assert res.expanded == ref["expanded"] and res.solutions == ref["solutions"]
    elif not res.split:
        # a single wavefront searched the whole tree: same DFS order as the oracle
        assert [list(m) for m in res.witness()] == ref["witness"]
        assert res.expanded == ref["expanded"]


def test_config1_iddfs_distances(native, exp):
    rows = exp["config1_iddfs"]
    for ms, col in ((native.MOVES_ALL, 1), (native.MOVES_REF, 2)):
        reqs = [native.linear_request(r[0], 10, 9, ms, native.FIRST, True, lex_order=True)
                for r in rows]
        run = native.batch(reqs)
        for q, r in enumerate(rows):
            assert run.results[q].found == 1 and run.results[q].depth == r[col]


def test_config2_first_witnesses(O, native, exp):
    ns, ks, perms = native.gen_products(1, 12, 16, 4, 9, 600)
    reqs = [native.linear_request(int(perms[q]), int(ns[q]), int(ks[q]), lex_order=True)
            for q in range(600)]
    run = native.batch(reqs)
    for row in exp["config2_first"]:
        q = row["q"]
        assert_matches(O, native, reqs[q], run.results[q], row)


def test_config3_depth8_and_nonexistence_proofs(O, native, exp):
    with open(os.path.join(GOLDEN, "instances.json")) as f:
        inst = json.load(f)
    _, _, p3 = native.gen_products(2, 16, 16, 9, 9, 1999)
    perms = list(p3) + [inst["config3"]["reverse"]]
    reqs = [native.linear_request(int(perms[q]), 16, 8, lex_order=True) for q in range(150)]
    run = native.batch(reqs)
    for row in exp["config3_first"]:
        assert_matches(O, native, reqs[row["q"]], run.results[row["q"]], row)
    # complete depth-8 trees: node counts are exact however the work was shared
    proofs = exp["config3_nonexistence"]
    reqs = [native.linear_request(p["perm"], 16, 8, lex_order=True) for p in proofs]
    for repeat in range(2):
        run = native.batch(reqs)
        for q, p in enumerate(proofs):
            r = run.results[q]
            assert r.found == 0
            assert r.expanded == p["expanded"] and r.solutions == 0


def test_count_all_fixtures(O, native, exp):
    reqs = []
    for row in exp["count_all"]:
        base = native.linear_request(row["perm"], row["n"], 9)
        beta, omega = pair(base)
        reqs.append(native.make_request(beta, omega, row["odd"], row["rate"], row["D"], True,
                                        row["moveset"], native.COUNT_ALL))
    run = native.batch(reqs)
    for q, row in enumerate(exp["count_all"]):
        assert_matches(O, native, reqs[q], run.results[q], row, exact=True)


def test_known_answers_every_path(O, native, exp):
    kap = exp["known_answers"]
    reqs = [native.make_request(k["beta"], k["omega"], k["initial_even"], k["min_rate"],
                                k["max_moves"]) for k in kap]
    batch = native.batch(reqs)
    native.init_slots(1, [0], 32 << 20)
    try:
        for q, k in enumerate(kap):
            pos = native.search_positions(0, k["beta"], k["omega"], k["initial_even"],
                                          k["min_rate"], k["max_moves"], True)
            assert bool(pos) == bool(k["found"]) == bool(batch.results[q].found)
            if pos:
                sym = native.positions_to_symbols(k["beta"], pos)
                assert O.validate(k["beta"], k["omega"], sym, k["max_moves"]) == 0
                if "witness_symbols" in k:
                    assert [list(s) for s in sym] == k["witness_symbols"]
    finally:
        native.shutdown()


def test_edge_cases(native):
    native.init_slots(1, [0], 32 << 20)
    try:
        # n < 3: no triple, empty result (sorting_search.cu:826-827)
        assert native.search_positions(0, [0, 1], [1, 0], 0, 1.0, 1) == []
        assert native.search_positions(0, [], [], 0, 1.0, 1) == []
        # already sorted: no move passes the reference filter
        assert native.search_positions(0, list(range(17)), list(range(17)), 17, 0.5, 9) == []
        # IDDFS on a sorted input: distance 0
        r = native.search_ex(0, native.make_request(list(range(5)), list(range(5)), 5, 0, 9,
                                                    iddfs=True))
        assert r.found == 1 and r.depth == 0
        # maximum sizes: n = 17 symbols, depth 9, the reverse permutation of 16
        rev = native.linear_request(sum(x << (4 * (15 - x)) for x in range(16)), 16, 9)
        pos = native.search_positions(0, *pair(rev), rev.initial_even, rev.min_rate, 9)
        assert len(pos) == 9
        assert native.search_positions(0, *pair(rev), rev.initial_even, rev.min_rate, 8) == []
        # invalid inputs
        for bad in (([0, 0, 2], [1, 2, 0]), ([0, 1, 2], [1, 1, 0]), ([0, 1, 3], [1, 2, 0])):
            with pytest.raises(Td16Error) as e:
                native.search_positions(0, bad[0], bad[1], 1, 2.0, 1)
            assert e.value.status == _lib.INVALID_ARGUMENT
        for mm in (0, 10):
            with pytest.raises(Td16Error):
                native.search_positions(0, [0, 1, 2], [1, 2, 0], 1, 2.0, mm)
        with pytest.raises(Td16Error) as e:
            native.search_positions(1, [0, 1, 2], [1, 2, 0], 1, 2.0, 1)
        assert e.value.status == _lib.INVALID_ARGUMENT  # slot out of range
    finally:
        native.shutdown()
    # budget x slots beyond free HBM (slot_manager.cu:87-106)
    with pytest.raises(Td16Error) as e:
        native.init_slots(2, [0], 1 << 50)
    assert e.value.status == _lib.OUT_OF_MEMORY
    with pytest.raises(Td16Error) as e:
        native.init_slots(1, [0, 0], 1 << 20)
    assert e.value.status == _lib.INVALID_ARGUMENT
    with pytest.raises(Td16Error) as e:
        native.init_slots(1, [7], 1 << 20)
    assert e.value.status == _lib.INVALID_ARGUMENT


def test_non_full_sorting_against_oracle(O, native):
    ns, ks, perms = native.gen_products(21, 8, 14, 3, 7, 120)
    reqs = []
    for q in range(120):
        base = native.linear_request(int(perms[q]), int(ns[q]), int(ks[q]))
        beta, omega = pair(base)
        reqs.append(native.make_request(beta, omega, base.initial_even, 1.25, int(ks[q]), False))
    run = native.batch(reqs)
    for q, req in enumerate(reqs):
        beta, omega = pair(req)
        ref = O.search(beta, omega, req.initial_even, req.min_rate, req.max_moves, False,
                       greedy=True)
        r = run.results[q]
        assert r.found == ref.found
        if not r.split:
            assert r.witness() == ref.witness() and r.expanded == ref.expanded


def test_dispatcher_smoke_like_reference():
    # ET/base/GPUSortingSearchSmokeTest.java:17-44: init({0}, 2 slots, 32 MiB),
    # 32 concurrent searches of the 3-cycle, each returns [[0, 1, 2]] (symbols)
    GPUSortingSearch.init([0], 2, 32 << 20)
    try:
        out = [None] * 32

        def one(i):
            out[i] = GPUSortingSearch.get_instance().search([0, 1, 2], [1, 2, 0], 1, 2.0, 1, True)

        ths = [threading.Thread(target=one, args=(i,)) for i in range(32)]
        for t in ths:
            t.start()
        for t in ths:
            t.join(60)
        assert all(o == [(0, 1, 2)] for o in out)
        inst = GPUSortingSearch.get_instance()
        assert inst.completed == 32 and inst.failed == 0
        # the witness facade: search -> replay -> certificate sink
        sink = MemoryCertificateSink()
        WitnessSearch.install_writer(sink)
        try:
            moves = WitnessSearch.search([0, 1, 2], [1, 2, 0], 2.0, 1)
            assert moves == [(0, 1, 2)]
            assert sink.find([1, 2, 0]) == [(0, 1, 2)]
        finally:
            WitnessSearch.uninstall_writer()
    finally:
        GPUSortingSearch.reset()


def replays(O, native, req, res):
    beta, omega = pair(req)
    sym = native.positions_to_symbols(beta, res.witness())
    return O.validate(beta, omega, sym, req.max_moves) == 0


def test_config1_full_size_distances(O, native, exp):
    """All 10,000 uniform n = 10 permutations (seed 0), exact distance by
    iterative deepening in both move sets, in one batched launch each: every
    distance equals the oracle's committed one, every witness replays."""
    perms = native.gen_uniform(0, 10, 10000)
    for ms, key in ((native.MOVES_ALL, "all"), (native.MOVES_REF, "ref")):
        want = exp["config1_distances"][key]
        reqs = [native.linear_request(int(p), 10, 9, ms, native.FIRST, True) for p in perms]
        run = native.batch(reqs)
        for q, req in enumerate(reqs):
            r = run.results[q]
            assert r.found == 1 and r.depth == int(want[q]), (key, q, r.depth, want[q])
            assert replays(O, native, req, r), (key, q)


def test_config1_plain_one_instance_per_launch(O, native, exp):
    """Configuration 1 as specified, the plain method: each of the 10,000
    permutations in its own launch through td16_search_ex (iterative
    deepening over all 165 triples, and again under the reference filter),
    seven slots each driven by its own thread as the reference's dispatcher
    drives them.  Distances equal the committed oracle distances."""
    perms = native.gen_uniform(0, 10, 10000)
    slots = 7
    native.init_slots(slots, [0], 64 << 20)
    try:
        for ms, key in ((native.MOVES_ALL, "all"), (native.MOVES_REF, "ref")):
            want = exp["config1_distances"][key]
            reqs = [native.linear_request(int(p), 10, 9, ms, native.FIRST, True) for p in perms]
            out = [None] * len(reqs)
            errors = []

            def worker(s):
                try:
                    for q in range(s, len(reqs), slots):
                        out[q] = native.search_ex(s, reqs[q])
                except Exception as e:  # surfaced below
                    errors.append(e)

            ths = [threading.Thread(target=worker, args=(s,)) for s in range(slots)]
            for t in ths:
                t.start()
            for t in ths:
                t.join(120)
            assert not errors, errors[:1]
            for q, req in enumerate(reqs):
                r = out[q]
                assert r.found == 1 and r.depth == int(want[q]), (key, q)
                if q % 10 == 0:
                    assert replays(O, native, req, r), (key, q)
    finally:
        native.shutdown()


def test_config3_every_instance_found_or_proved(O, native, exp):
    """All 2,000 n = 16 instances (1,999 products of 9 transpositions and the
    reverse permutation) at move bound 8 under the reference filter, in the
    kernel's default order: found / not-found equals the oracle's for every
    instance, every witness replays, and the non-existence proofs (complete
    trees) have the oracle's node counts exactly."""
    with open(os.path.join(GOLDEN, "instances.json")) as f:
        inst = json.load(f)
    _, _, p3 = native.gen_products(2, 16, 16, 9, 9, 1999)
    perms = list(p3) + [inst["config3"]["reverse"]]
    reqs = [native.linear_request(int(p), 16, 8) for p in perms]
    run = native.batch(reqs)
    want = exp["config3_found"]
    proofs = {p["q"]: p for p in exp["config3_nonexistence"]}
    for q, req in enumerate(reqs):
        r = run.results[q]
        assert r.found == int(want[q]), q
        if r.found:
            assert replays(O, native, req, r), q
        else:
            assert (r.expanded, r.solutions) == (proofs[q]["expanded"], 0), q


def test_config3_exhaustive_counts(native, exp):
    """Exhaustive depth-8 searches (COUNT_ALL) of the zero-slack n = 16
    instances whose trees the CPU finishes: node and solution counts equal
    the oracle's bit for bit, under the full move set and under the
    reference filter (zero slack leaves only moves that gain two odd
    cycles, so both trees are the same)."""
    rows = exp["config3_count_all"]["rows"]
    assert len(rows) >= 2
    for ms in (native.MOVES_ALL, native.MOVES_REF):
        reqs = [native.linear_request(r["perm"], 16, 8, ms, native.COUNT_ALL) for r in rows]
        run = native.batch(reqs, waves_per_cu=20)
        for q, row in enumerate(rows):
            r = run.results[q]
            assert (r.expanded, r.solutions) == (row["expanded"], row["solutions"]), (ms, row["q"])
