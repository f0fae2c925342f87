"""GPU parity: the HIP search (through the C ABI) against the CPU oracle.

Bars: found / not-found, witness replay validity, and -- for searches that a
single wavefront ran from start to finish (result.split == 0) or that were
exhaustive -- bit-exact witnesses and node counts.  FIRST searches are
compared with the oracle in the same visiting order: greedy by default,
lexicographic (the reference's) when the request sets lex_order.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def slots(native):
    native.init_slots(2, [0], 64 << 20)
    yield
    native.shutdown()


def pair_of(req):
    return list(req.beta[:req.n]), list(req.omega[:req.n])


def check_against_oracle(O, native, req, res, exact_counts=False):
    beta, omega = pair_of(req)
    greedy = req.mode == 0 and not req.lex_order
    if req.iddfs:
        d, ref = O.iddfs(beta, omega, req.moveset, req.max_moves, greedy=greedy)
        assert res.found == (d >= 0)
        if d >= 0:
            assert res.depth == d
    else:
        ref = O.search(beta, omega, req.initial_even, req.min_rate, req.max_moves,
                       bool(req.full_sorting), req.moveset, req.mode, greedy=greedy)
        assert res.found == ref.found
    if res.found and res.depth > 0:
        sym = native.positions_to_symbols(beta, res.witness())
        assert O.validate(beta, omega, sym, req.max_moves) == 0
    if req.mode == 1 or not res.found:
        assert res.expanded == ref.expanded
        assert res.solutions == ref.solutions
    elif not res.split:
        assert res.witness() == ref.witness()
        assert res.expanded == ref.expanded


def test_reference_smoke_vector(native, slots):
    # GPUSortingSearchSmokeTest.java:26-35
    moves = native.search_positions(0, [0, 1, 2], [1, 2, 0], 1, 2.0, 1, True)
    assert moves == [(0, 1, 2)]


def test_plain_against_oracle(O, native, slots):
    ns, ks, perms = native.gen_products(11, 8, 16, 3, 9, 48)
    for q in range(len(ns)):
        req = native.linear_request(int(perms[q]), int(ns[q]), int(ks[q]))
        res = native.search_ex(q % 2, req)
        check_against_oracle(O, native, req, res)


def test_batch_config2_sample(O, native):
    ns, ks, perms = native.gen_products(1, 12, 16, 4, 9, 400)
    reqs = [native.linear_request(int(perms[q]), int(ns[q]), int(ks[q])) for q in range(400)]
    run = native.batch(reqs)
    for q in range(400):
        check_against_oracle(O, native, reqs[q], run.results[q])


def bounded_request(native, perm, n, slack, moveset, mode):
    """COUNT_ALL-sized request: move bound = odd-cycle bound + slack."""
    base = native.linear_request(int(perm), int(n), 9, moveset, mode)
    lb = (base.n - base.initial_even) // 2
    D = min(9, lb + slack)
    beta, omega = pair_of(base)
    return native.make_request(beta, omega, base.initial_even, 2.0 / (D / float(lb)), D, True,
                               moveset, mode)


def test_batch_count_all(O, native):
    ns, ks, perms = native.gen_products(5, 8, 11, 4, 6, 64)
    reqs = []
    for q in range(64):
        for ms in (native.MOVES_REF, native.MOVES_ALL):
            reqs.append(bounded_request(native, perms[q], ns[q], 0, ms, native.COUNT_ALL))
    ns, ks, perms = native.gen_products(6, 6, 8, 3, 5, 32)
    for q in range(32):
        reqs.append(bounded_request(native, perms[q], ns[q], 1, native.MOVES_REF, native.COUNT_ALL))
    run = native.batch(reqs)
    for q, req in enumerate(reqs):
        check_against_oracle(O, native, req, run.results[q], exact_counts=True)


def test_batch_iddfs_all_n7(O, native):
    _, dist, _ = O.bfs(7)
    reqs = []
    for r in range(1, math.factorial(7)):
        p = O.perm_unrank(r, 7)
        reqs.append(native.linear_request(O.pack(p), 7, 9, native.MOVES_ALL, native.FIRST, True))
    run = native.batch(reqs)
    for q in range(len(reqs)):
        assert run.results[q].found == 1
        assert run.results[q].depth == dist[q + 1]


def test_both_orders_against_oracle(O, native, slots):
    """Searches of a few hundred nodes in both visiting orders: each matches
    the oracle in its own order (exactly when one wavefront ran it), and the
    greedy order needs no more nodes than the lexicographic one overall."""
    ns, ks, perms = native.gen_products(1, 12, 16, 4, 9, 600)
    picked = []
    for q in range(len(ns)):
        req = native.linear_request(int(perms[q]), int(ns[q]), int(ks[q]), lex_order=True)
        beta, omega = pair_of(req)
        ref = O.search(beta, omega, req.initial_even, req.min_rate, req.max_moves,
                       bool(req.full_sorting), req.moveset, req.mode)
        if ref.found and 40 <= ref.expanded <= 20000:
            picked.append(req)
        if len(picked) == 12:
            break
    assert picked
    nodes = {0: 0, 1: 0}
    for req in picked:
        for lex in (1, 0):
            req.lex_order = lex
            res = native.search_ex(lex, req)
            check_against_oracle(O, native, req, res)
            nodes[lex] += res.expanded
    assert nodes[0] <= nodes[1]


@pytest.mark.parametrize("count", [1, 2, 63, 64, 65, 129, 1000])
def test_batch_sizes_around_feed_shards(O, native, count):
    """Batches smaller than, equal to and just above the 64 feed shards (empty
    and short shards, look-ahead claims near shard ends), mixed with requests
    the host answers itself (n < 3, sorted IDDFS): every result belongs to its
    own request."""
    ns, ks, perms = native.gen_products(31 + count, 8, 14, 3, 7, count)
    reqs = []
    for q in range(count):
        if q % 17 == 5:
            reqs.append(native.make_request([0, 1], [1, 0], 0, 1.0, 3))       # n < 3
        elif q % 17 == 11:
            reqs.append(native.make_request(list(range(6)), list(range(6)), 6, 0.0, 5,
                                            iddfs=True))                      # sorted
        else:
            reqs.append(native.linear_request(int(perms[q]), int(ns[q]), int(ks[q])))
    run = native.batch(reqs)
    for q, req in enumerate(reqs):
        r = run.results[q]
        assert r.status == 0
        if q % 17 == 5:
            assert r.found == 0 and r.depth == 0
        elif q % 17 == 11:
            assert r.found == 1 and r.depth == 0
        else:
            check_against_oracle(O, native, req, r)


def test_emitter_stream_certified_and_checked(O, native):
    """An emitter-shaped stream (n = 16, depth bound 9, one instance in ten
    reverse-like): the first 1,000 against the oracle (greedy order), and every
    witness of all 20,000 certified (replay + codec records) and checked again
    from its stored record by the batched checker."""
    count = 20000
    ns, ks, perms = native.gen_emitter_stream(5, count)
    reqs = [native.linear_request(int(perms[q]), int(ns[q]), int(ks[q])) for q in range(count)]
    run = native.batch(reqs)
    sample = 1000
    ref = O.search_batch(ns[:sample], perms[:sample], ks[:sample], O.MOVES_REF, O.FIRST,
                         threads=8, greedy=True)
    for q in range(sample):
        r = run.results[q]
        assert r.status == 0 and r.found == ref[q].found
        if not r.split:
            assert r.witness() == ref[q].witness() and r.expanded == ref[q].expanded
    found = np.array([run.results[q].status == 0 and run.results[q].found == 1
                      for q in range(count)])
    certs = native.certify_batch(reqs, run.results)
    assert np.array_equal(certs.status == 0, found)
    assert np.all(certs.status[~found] == 1)
    betas, omegas, nn, mm = native.obligations((type(reqs[0]) * count)(*reqs))
    st = native.check_batch(betas, omegas, nn, mm, certs.values, certs.value_len)
    assert np.array_equal(st == 0, found) and np.all(st[~found] == 1)
    # a record is the codec of the witness's symbols under the request's omega key
    for q in range(0, count, 997):
        if found[q]:
            beta, omega = pair_of(reqs[q])
            sym = native.positions_to_symbols(beta, run.results[q].witness())
            assert certs.value(q) == O.codec_value(sym)
            assert certs.key(q, reqs[q].n) == O.codec_key(omega)


def test_unoriented_stage_partitions_all_solved(O, native):
    """The emitter's unoriented stage on every partition with at most 20,000
    classes (32,880 searches of n = 17): each class has a witness of at most 9
    moves (emitter.log reports residual=0 for every partition), each witness
    replays and certifies, and a sample matches the oracle in the same order."""
    from td16_amd import unoriented
    cases = [(p, t) for p, t in unoriented.STAGE if native.bracelet_count(p) <= 20000]
    sb = unoriented.stage_requests(cases)
    assert sb.classes == [native.bracelet_count(p) for p, _ in cases]
    count = len(sb.requests)
    assert count == 32880
    run = native.batch(sb.requests)
    f = native.result_fields(run.results)
    assert np.all(f["status"] == 0) and np.all(f["found"] == 1)
    assert np.all((f["depth"] >= 1) & (f["depth"] <= 9))
    certs = native.certify_batch(sb.requests, run.results)
    assert np.all(certs.status == 0)
    st = native.check_batch(*native.obligations(sb.requests), certs.values, certs.value_len)
    assert np.all(st == 0)
    for q in range(0, count, 97):
        req = sb.requests[q]
        beta, omega = pair_of(req)
        ref = O.search(beta, omega, req.initial_even, req.min_rate, req.max_moves, True,
                       greedy=True)
        r = run.results[q]
        assert ref.found == 1
        if not r.split:
            assert r.witness() == ref.witness() and r.expanded == ref.expanded


# Settings of the three sharing knobs (wavefronts per CU, donation depth,
# steal granularity), from the sparsest grid to the densest.
KNOB_POINTS = [dict(waves_per_cu=4, split_depth=2, donate_max=4, check_every=1),
               dict(waves_per_cu=8, split_depth=7, donate_max=64, check_every=32),
               dict(waves_per_cu=16, split_depth=3, donate_max=16, check_every=8),
               dict(waves_per_cu=12, split_depth=5, donate_max=32, check_every=1024),
               dict(waves_per_cu=20, split_depth=4, donate_max=8, check_every=16),
               # the 6-waves-per-SIMD build of the kernel
               dict(waves_per_cu=24, split_depth=6, donate_max=16, check_every=16)]
KNOBS = ("waves_per_cu", "split_depth", "donate_max", "check_every")


def knob_requests(native):
    """Exhaustive searches in both move sets plus a slice of the batched mix."""
    ns, ks, perms = native.gen_products(5, 8, 11, 4, 6, 48)
    reqs = [bounded_request(native, perms[q], ns[q], 0, ms, native.COUNT_ALL)
            for q in range(48) for ms in (native.MOVES_REF, native.MOVES_ALL)]
    ns, ks, perms = native.gen_products(1, 12, 16, 4, 9, 600)
    reqs += [native.linear_request(int(perms[q]), int(ns[q]), int(ks[q])) for q in range(600)]
    return reqs


def test_sharing_knobs_never_change_answers(O, native):
    """The knobs decide how work is shared, not what is found: exhaustive
    counts equal the oracle's at every setting, and found / not-found never
    moves."""
    reqs = knob_requests(native)
    runs = [native.batch(reqs, **k) for k in KNOB_POINTS]
    assert sum(r.stats.donations for r in runs) > 0
    for q, req in enumerate(reqs):
        for run in runs:
            check_against_oracle(O, native, req, run.results[q])


def test_autotune_picks_a_timed_point_and_keeps_answers(O, native):
    reqs = knob_requests(native)
    t = native.autotune(reqs, repeats=1)
    assert tuple(t.points[0][k] for k in KNOBS) == (16, 5, 16, 16)  # the library defaults
    # the first pass times every value of every knob once
    assert len(t.points) >= 5 + 5 + 4 + 2
    chosen = [p for p in t.points if all(p[k] == t.best[k] for k in KNOBS)]
    assert len(chosen) == 1
    assert chosen[0]["kernel_ms"] <= t.points[0]["kernel_ms"]
    # wavefronts per CU as asked: the grid scales with the knob
    assert len({p["waves"] // p["waves_per_cu"] for p in t.points}) == 1
    for q, req in enumerate(reqs):
        check_against_oracle(O, native, req, t.results[q])
    run = native.batch(reqs, **t.best)
    for q, req in enumerate(reqs):
        check_against_oracle(O, native, req, run.results[q])


def mixed_batch(native, seed, count):
    """Batched-mix products in both orders, with n = 10 exact distances, a
    sorted input and two-symbol requests mixed in."""
    ns, ks, perms = native.gen_products(seed, 12, 16, 4, 9, count)
    reqs = [native.linear_request(int(perms[q]), int(ns[q]), int(ks[q]), lex_order=(q % 4 == 0))
            for q in range(count)]
    for p in native.gen_uniform(seed, 10, count // 10):
        reqs.append(native.linear_request(int(p), 10, 9, native.MOVES_REF, native.FIRST, True))
    reqs.append(native.make_request(list(range(6)), list(range(6)), 6, 0.0, 5, iddfs=True))
    reqs.append(native.make_request([0, 1], [1, 0], 0, 1.0, 1))
    return reqs


def same_answers(native, reqs, a, b):
    both_unsplit = 0
    for q, req in enumerate(reqs):
        x, y = a[q], b[q]
        assert (x.status, x.found) == (y.status, y.found), q
        if req.iddfs:
            assert x.depth == y.depth, q
        if not x.split and not y.split:
            assert x.witness() == y.witness() and x.expanded == y.expanded, q
            both_unsplit += 1
    return both_unsplit


@pytest.mark.parametrize("chunk", [1000, 7777])
def test_stream_matches_single_batches(O, native, chunk):
    """td16_batch_stream (double-buffered chunks) gives td16_batch's answers:
    chunk sizes that do and do not divide the stream, trivial requests in
    any chunk."""
    reqs = mixed_batch(native, 41, 20000)
    ref = native.batch(reqs)
    got = native.batch_stream(reqs, chunk)
    assert same_answers(native, reqs, ref.results, got.results) > len(reqs) // 2
    for q in range(0, len(reqs), 97):
        check_against_oracle(O, native, reqs[q], got.results[q])
    st = got.stats
    assert st.total_ms > 0 and st.kernel_ms > 0 and st.expanded > 0


def test_prepare_thread_counts_give_the_same_batch(native):
    """The host preparation is deterministic: one thread or sixteen, the
    launch sees the same requests in the same feed order."""
    reqs = mixed_batch(native, 43, 30000)
    a = native.batch(reqs, host_threads=1)
    b = native.batch(reqs, host_threads=16)
    assert same_answers(native, reqs, a.results, b.results) > len(reqs) // 2
    assert a.stats.prepare_ms > 0 and b.stats.decode_ms > 0


def test_counted_build_counts_and_keeps_answers(native):
    """The counted build (count_candidates) finds what the timed build finds
    and sums C(m,3) over the expanded nodes: exactly 1 for the reference's
    smoke vector (one node, three breakpoints), and between one and C(N,3)
    per expanded node on a mix."""
    smoke = native.make_request([0, 1, 2], [1, 2, 0], 1, 2.0, 1)
    run = native.batch([smoke], count_candidates=True)
    assert run.results[0].found == 1 and run.stats.candidates == 1
    reqs = mixed_batch(native, 47, 5000)
    timed = native.batch(reqs)
    counted = native.batch(reqs, count_candidates=True)
    assert same_answers(native, reqs, timed.results, counted.results) > len(reqs) // 2
    top = sum(counted.results[q].expanded * math.comb(reqs[q].n, 3) for q in range(len(reqs)))
    assert counted.stats.expanded <= counted.stats.candidates <= top
    assert timed.stats.candidates == 0
