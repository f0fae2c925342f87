"""Served loop (continuous batching into one persistent kernel) and the
hipGraph batched launch: every result is the same alone and in the mix."""
import threading
import time

import pytest

pytestmark = pytest.mark.gpu


def _reqs(native, seed, count, kmin=4, kmax=9):
    ns, ks, perms = native.gen_products(seed, 12, 16, kmin, kmax, count)
    return [native.linear_request(int(perms[q]), int(ns[q]), int(ks[q])) for q in range(count)]


def _pair(req):
    return list(req.beta[:req.n]), list(req.omega[:req.n])


def _same_answer(O, native, req, a, b):
    assert a.found == b.found
    for r in (a, b):
        if r.found:
            beta, omega = _pair(req)
            sym = native.positions_to_symbols(beta, r.witness())
            assert O.validate(beta, omega, sym, req.max_moves) == 0
    if not a.split and not b.split:
        assert a.witness() == b.witness()
        assert a.expanded == b.expanded


def test_served_mix_matches_alone(O, native):
    reqs = _reqs(native, 3, 300, 6, 9)
    alone = native.batch(reqs)
    native.serve_start(device=0, slots=32, queue_capacity=512, slot_bytes=16 << 20)
    try:
        tickets = [native.serve_submit(r) for r in reqs]
        mixed = [native.serve_wait(t) for t in tickets]
        st = native.serve_stats()
        assert st.completed >= len(reqs)
        assert st.p99_us >= st.p50_us > 0
        # the same request alone in the running loop
        for q in (0, 7, 99):
            solo = native.serve_search(reqs[q])
            _same_answer(O, native, reqs[q], solo, mixed[q])
    finally:
        native.serve_stop()
    for q, req in enumerate(reqs):
        _same_answer(O, native, req, alone.results[q], mixed[q])
        ref = O.search(*_pair(req), req.initial_even, req.min_rate, req.max_moves, True)
        assert mixed[q].found == ref.found


def test_served_concurrent_producers(O, native):
    reqs = _reqs(native, 4, 240)
    native.serve_start(device=0, slots=8, queue_capacity=16, slot_bytes=16 << 20)
    out = [None] * len(reqs)
    try:
        def worker(lo):
            for q in range(lo, len(reqs), 6):
                out[q] = native.serve_search(reqs[q])
        ths = [threading.Thread(target=worker, args=(i,)) for i in range(6)]
        for t in ths:
            t.start()
        for t in ths:
            t.join(120)
    finally:
        native.serve_stop()
    for q, req in enumerate(reqs):
        ref = O.search(*_pair(req), req.initial_even, req.min_rate, req.max_moves, True)
        assert out[q] is not None and out[q].found == ref.found


def test_graph_replays_match_eager(native):
    reqs = _reqs(native, 9, 500)
    eager = native.batch(reqs)
    for rep in range(3):
        g = native.batch(reqs, use_graph=True, repeats=3)
        # eager work and a sync between replays
        native.batch(reqs[:50])
        for q in range(len(reqs)):
            assert g.results[q].found == eager.results[q].found
            if not g.results[q].split and not eager.results[q].split:
                assert g.results[q].witness() == eager.results[q].witness()


def _golden():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "expected.json")
    with open(path) as f:
        return json.load(f)


def test_served_budget_keeps_exhaustive_counts(native):
    """A per-slot budget of eight queued subtrees (the loop allocates slots x
    budget and each admitted request keeps at most its budget queued): the
    exhaustive searches still count every node, since a full budget only
    keeps work with the wavefronts that hold it.  COUNT_ALL fixtures and the
    n = 16 depth-8 non-existence proofs against the oracle's counts."""
    exp = _golden()
    reqs, want = [], []
    for row in exp["count_all"][:60]:
        base = native.linear_request(row["perm"], row["n"], 9)
        reqs.append(native.make_request(list(base.beta[:base.n]), list(base.omega[:base.n]),
                                        row["odd"], row["rate"], row["D"], True,
                                        row["moveset"], native.COUNT_ALL))
        want.append((row["expanded"], row["solutions"]))
    for p in exp["config3_nonexistence"]:
        reqs.append(native.linear_request(p["perm"], 16, 8))
        want.append((p["expanded"], 0))
    item = 48 + 8
    native.serve_start(device=0, slots=8, queue_capacity=128, slot_bytes=8 * item)
    try:
        tickets = [native.serve_submit(r) for r in reqs]
        got = [native.serve_wait(t) for t in tickets]
    finally:
        native.serve_stop()
    for q, r in enumerate(got):
        assert (r.expanded, r.solutions) == want[q], q


def test_served_budget_checked_against_free_memory(native):
    from td16_amd import _lib
    with pytest.raises(_lib.Td16Error) as e:
        native.serve_start(device=0, slots=64, slot_bytes=1 << 40)
    assert e.value.status == _lib.OUT_OF_MEMORY
    with pytest.raises(_lib.Td16Error) as e:
        native.serve_start(device=0, slots=8, slot_bytes=16)
    assert e.value.status == _lib.INVALID_ARGUMENT


def test_served_closed_loop_matches_batch(O, native):
    """Many producers submitting and waiting with no pacing: every answer
    equals the batched launch's, every witness replays."""
    reqs = _reqs(native, 12, 6000, 6, 9)
    alone = native.batch(reqs)
    native.serve_start(device=0, slots=64, queue_capacity=512, slot_bytes=16 << 20)
    try:
        out, elapsed_us, st = native.serve_closed_loop(reqs, 32)
    finally:
        native.serve_stop()
    assert st.completed >= len(reqs) and st.p99_us >= st.p50_us > 0
    for q in range(len(reqs)):
        _same_answer(O, native, reqs[q], alone.results[q], out[q])
    assert len(reqs) / (elapsed_us * 1e-6) > 1000


def test_stop_while_producers_submit(native):
    """A stop that races with producer threads: each producer's calls either
    return a result or fail with NOT_INITIALIZED once the stop has begun;
    nothing hangs and nothing else fails."""
    from td16_amd import _lib
    reqs = _reqs(native, 13, 64, 5, 8)
    native.serve_start(device=0, slots=8, queue_capacity=4, slot_bytes=16 << 20)
    done, errors, stopped = [0] * 8, [], [0] * 8

    def producer(p):
        q = p
        while True:
            try:
                r = native.serve_search(reqs[q % len(reqs)])
                assert r.status == 0
                done[p] += 1
            except _lib.Td16Error as e:
                if e.status == _lib.NOT_INITIALIZED:
                    stopped[p] = 1
                    return
                errors.append(e)
                return
            q += 8

    ths = [threading.Thread(target=producer, args=(p,)) for p in range(8)]
    for t in ths:
        t.start()
    time.sleep(0.2)
    native.serve_stop()
    for t in ths:
        t.join(30)
    assert not any(t.is_alive() for t in ths)
    assert not errors, errors[:1]
    assert all(stopped) and sum(done) > 0


def test_wait_deadline_reports_and_the_ticket_survives(native):
    """With a 1 ms deadline, waiting on a search that needs millions of
    nodes returns TD16_DEVICE_ERROR; the ticket stays valid and a later wait
    collects the exact result."""
    from td16_amd import _lib
    proof = max(_golden()["config3_nonexistence"], key=lambda p: p["expanded"])
    req = native.linear_request(proof["perm"], 16, 8)
    native.serve_start(device=0, slots=4, queue_capacity=8, slot_bytes=16 << 20, deadline_ms=1)
    try:
        t = native.serve_submit(req)
        with pytest.raises(_lib.Td16Error) as e:
            native.serve_wait(t)
        assert e.value.status == _lib.DEVICE_ERROR and "deadline" in str(e.value)
        res = None
        for _ in range(5000):
            try:
                res = native.serve_wait(t)
                break
            except _lib.Td16Error as e2:
                assert e2.status == _lib.DEVICE_ERROR
        assert res is not None and res.found == 0 and res.expanded == proof["expanded"]
    finally:
        native.serve_stop()
