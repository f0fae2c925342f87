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
