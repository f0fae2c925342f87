"""td16_batch_noalloc (no device allocation; one static device array per GPU)
against td16_batch, through the C ABI: two calls in a row, the second with a
smaller batch, so the second call runs on the array the first one left
behind.  Same answers to the bar of every other GPU parity test: found /
not-found per request, exact iterative-deepening distances, exact COUNT_ALL
node and solution counts, and for requests searched by one wavefront in both
calls the same witness and node count; every witness replays in the oracle's
checker."""
import json
import os

import pytest

from td16_amd import _lib

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def requests(native, count, seed):
    """Mixed batch: products of the batched mix (greedy and lexicographic
    order), n = 10 exact distances in both move sets, and the COUNT_ALL
    fixtures."""
    with open(os.path.join(GOLDEN, "expected.json")) as f:
        exp = json.load(f)
    ns, ks, perms = native.gen_products(seed, 12, 16, 4, 9, count)
    reqs = [native.linear_request(int(perms[q]), int(ns[q]), int(ks[q]), lex_order=(q % 3 == 0))
            for q in range(count)]
    for r in exp["config1_iddfs"][: count // 20]:
        for ms in (native.MOVES_ALL, native.MOVES_REF):
            reqs.append(native.linear_request(r[0], 10, 9, ms, native.FIRST, True))
    for row in exp["count_all"][: count // 40]:
        base = native.linear_request(row["perm"], row["n"], 9)
        beta, omega = list(base.beta[:base.n]), list(base.omega[:base.n])
        reqs.append(native.make_request(beta, omega, row["odd"], row["rate"], row["D"], True,
                                        row["moveset"], native.COUNT_ALL))
    return reqs


def check_same(O, native, reqs, ref, got):
    shared = 0
    for q, req in enumerate(reqs):
        a, b = ref.results[q], got.results[q]
        assert a.status == b.status == 0, q
        assert a.found == b.found, q
        if req.iddfs:
            assert a.depth == b.depth, q
        if req.mode == native.COUNT_ALL:
            assert (a.expanded, a.solutions) == (b.expanded, b.solutions), q
        elif not a.split and not b.split:
            assert a.witness() == b.witness() and a.expanded == b.expanded, q
            shared += 1
        if b.found and b.depth:
            beta, omega = list(req.beta[:req.n]), list(req.omega[:req.n])
            sym = native.positions_to_symbols(beta, b.witness())
            assert O.validate(beta, omega, sym, req.max_moves) == 0, q
    assert shared > len(reqs) // 2  # most requests are compared witness for witness


def test_noalloc_matches_batch_on_two_calls(O, native):
    for count, seed in ((8000, 31), (1200, 32)):
        reqs = requests(native, count, seed)
        print("noalloc case: %d requests, seed %d" % (len(reqs), seed), flush=True)
        assert len(reqs) <= _lib.NOALLOC_MAX_REQUESTS
        ref = native.batch(reqs)
        got = native.batch_noalloc(reqs)
        check_same(O, native, reqs, ref, got)
        assert got.stats.expanded > 0 and got.stats.waves == ref.stats.waves
