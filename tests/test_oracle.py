"""Pins the CPU oracle (oracle/td16_oracle.c) against facts that do not come
from the oracle itself: brute-force recounts, BFS over S_n, closed forms, the
reference's own fixtures and logged outcomes."""
import itertools
import json
import math
import os
import random

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def recount_after(omega, a, b, c):
    """omega o (a b c)^-1 by plain composition: x -> omega[inv(x)]."""
    inv = list(range(len(omega)))
    inv[a], inv[b], inv[c] = c, a, b
    return [omega[inv[x]] for x in range(len(omega))]


def cycle_sets(omega):
    seen, out = set(), {}
    for s in range(len(omega)):
        if s in seen:
            continue
        x, cyc = s, []
        while x not in seen:
            seen.add(x)
            cyc.append(x)
            x = omega[x]
        for y in cyc:
            out[y] = s
    return out


def odd_cycles_py(omega):
    lab = cycle_sets(omega)
    sizes = {}
    for v in lab.values():
        sizes[v] = sizes.get(v, 0) + 1
    return sum(1 for s in sizes.values() if s & 1)


def test_count_odd_cycles_closed_forms(O):
    for n in range(1, 18):
        assert O.count_odd_cycles(list(range(n))) == n               # identity
        assert O.count_odd_cycles([(x + 1) % n for x in range(n)]) == (n & 1)  # n-cycle
    rng = random.Random(1)
    for _ in range(500):
        n = rng.randint(1, 17)
        p = list(range(n))
        rng.shuffle(p)
        c = O.count_odd_cycles(p)
        assert c == odd_cycles_py(p)
        assert (c - n) % 2 == 0  # sum of lengths = n


@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_delta_equals_recount_exhaustive(O, n):
    # sorting_search.cu:240-332 restated; every ordered triple of distinct
    # symbols of every permutation of size n
    for p in itertools.permutations(range(n)):
        omega = list(p)
        base = odd_cycles_py(omega)
        lab = cycle_sets(omega)
        for a, b, c in itertools.permutations(range(n), 3):
            d = O.delta(omega, a, b, c)
            three = lab[a] != lab[b] and lab[b] != lab[c] and lab[a] != lab[c]
            if three:
                assert d is None
            else:
                assert d == odd_cycles_py(recount_after(omega, a, b, c)) - base


def test_delta_equals_recount_sampled_n17(O):
    rng = random.Random(7)
    for _ in range(3000):
        omega = list(range(17))
        rng.shuffle(omega)
        a, b, c = rng.sample(range(17), 3)
        lab = cycle_sets(omega)
        d = O.delta(omega, a, b, c)
        if lab[a] != lab[b] and lab[b] != lab[c] and lab[a] != lab[c]:
            assert d is None
        else:
            assert d == odd_cycles_py(recount_after(omega, a, b, c)) - odd_cycles_py(omega)


# distance histograms of S_n under block transpositions (counts of
# permutations at distance 0, 1, 2, ...): the published values of the
# transposition distance distribution
HIST = {
    4: [1, 10, 12, 1],
    5: [1, 20, 68, 31],
    6: [1, 35, 259, 380, 45],
    7: [1, 56, 770, 2700, 1513],
    8: [1, 84, 1932, 13467, 22000, 2836],
}
# diameters from the reference's table, E/base/CommonOperations.java:178-192
DIAMETER = {2: 1, 3: 2, 4: 3, 5: 3, 6: 4, 7: 4, 8: 5}


def test_bfs_histograms_and_diameters(O):
    for n in range(2, 9):
        diam, dist, _ = O.bfs(n)
        assert diam == DIAMETER[n]
        if n in HIST:
            assert list(np.bincount(dist)) == HIST[n]
        # level 1 holds exactly the C(n+1,3) distinct block transpositions
        assert int(np.sum(dist == 1)) == math.comb(n + 1, 3)


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6])
def test_iddfs_equals_bfs_all_permutations(O, n):
    _, dist, _ = O.bfs(n)
    for r in range(math.factorial(n)):
        p = O.perm_unrank(r, n)
        beta, omega = O.linear_to_pair(O.pack(p), n)
        for ms in (O.MOVES_ALL, O.MOVES_REF):
            d, res = O.iddfs(beta, omega, ms, 9)
            assert d == dist[r], (p, ms)
            if d > 0:
                sym = O.positions_to_symbols(beta, res.witness())
                assert O.validate(beta, omega, sym, d) == 0


def test_iddfs_equals_bfs_n7_sample(O):
    _, dist, _ = O.bfs(7)
    rng = random.Random(3)
    for r in rng.sample(range(math.factorial(7)), 600):
        beta, omega = O.linear_to_pair(O.pack(O.perm_unrank(r, 7)), 7)
        assert O.iddfs(beta, omega, O.MOVES_ALL, 9)[0] == dist[r]
        assert O.iddfs(beta, omega, O.MOVES_REF, 9)[0] == dist[r]


@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_count_all_solutions_equal_shortest_paths(O, n):
    # with the bound at the exact distance, every sorting sequence of that
    # length is a shortest path of the Cayley graph and no prefix is pruned
    _, dist, paths = O.bfs(n, with_paths=True)
    for r in range(1, math.factorial(n)):
        beta, omega = O.linear_to_pair(O.pack(O.perm_unrank(r, n)), n)
        odd = O.count_odd_cycles(omega)
        sat, reach = O.rate_tables(n + 1, odd, 0.0, int(dist[r]), True)
        res = O.search_tables(beta, omega, int(dist[r]), sat, reach, O.MOVES_ALL, O.COUNT_ALL)
        assert res.solutions == paths[r]


def test_count_all_equals_naive_enumeration(O):
    rng = random.Random(11)
    for _ in range(60):
        n = rng.randint(3, 5)
        p = list(range(n))
        rng.shuffle(p)
        if p == sorted(p):
            continue
        beta, omega = O.linear_to_pair(O.pack(p), n)
        odd = O.count_odd_cycles(omega)
        lb = (n + 1 - odd) // 2
        D = min(4, lb + rng.randint(0, 2))
        if D < 1:
            continue
        for full in (True, False):
            rate = rng.choice([0.0, 0.5, 1.0, 2.0])
            sat, reach = O.rate_tables(n + 1, odd, rate, D, full)
            for ms in (O.MOVES_REF, O.MOVES_ALL):
                res = O.search_tables(beta, omega, D, sat, reach, ms, O.COUNT_ALL)
                e, s = O.count_naive(beta, omega, D, sat, reach, ms)
                assert (res.expanded, res.solutions) == (e, s)


def test_rate_tables_closed_forms(O):
    # fullSorting with the emitter's convention minRate = 2/(k/(float)t):
    # a witness is exactly "sorted within k" (E/UnorientedCycles.java:34)
    for N in range(3, 18):
        for odd in range(N % 2, N, 2):
            t = (N - odd) // 2
            for k in range(1, 10):
                rate = float(np.float32(2.0) / (np.float32(k) / np.float32(t)))
                sat, reach = O.rate_tables(N, odd, rate, k, True)
                for d in range(10):
                    assert sat[d] == ((1 << N) if 1 <= d <= k else 0), (N, odd, k, d)
                    # can_still_reach: d + ceil((N-e)/2) <= k
                    want = sum(1 << e for e in range(N + 1) if d + (N - e + 1) // 2 <= k)
                    assert reach[d] == want


def test_rate_tables_non_full_sorting(O):
    # sorting_search.cu:343-345, 357-370 by hand for N=9, ie=1, minRate 1.5,
    # maxDepth 4: required gain at max depth = floor(1.5*4) = 6
    sat, reach = O.rate_tables(9, 1, 1.5, 4, False)
    for d in range(1, 5):
        for e in range(10):
            g = e - 1
            assert bool(sat[d] >> e & 1) == (g > 0 and g / d >= 1.5)
    # depth 2, e = 3: gain 2, optimistic 2 + 2*2 = 6 >= 6; lb = 3, best
    # ratio (2 + 6) / (2 + 3) = 1.6 >= 1.5 -> reachable
    assert reach[2] >> 3 & 1
    # depth 3, e = 3: optimistic 2 + 2 = 4 < 6 -> pruned
    assert not reach[3] >> 3 & 1
    # NaN minRate: no witness ever, and floor(NaN*4) converts to 0
    sat, reach = O.rate_tables(9, 1, float("nan"), 4, False)
    assert all(s == 0 for s in sat)


def test_reference_smoke_vector(O):
    # ET/base/GPUSortingSearchSmokeTest.java:26-35
    r = O.search([0, 1, 2], [1, 2, 0], 1, 2.0, 1, True)
    assert r.found == 1 and r.witness() == [(0, 1, 2)]
    assert O.positions_to_symbols([0, 1, 2], r.witness()) == [(0, 1, 2)]


def test_reference_algebra_fixtures(O):
    # ET/permutation/PermutationGroupsTest.java:9-12: (0 4 1)(0 4 5 6 1 2 3) =
    # (0 1 2 3 4 5 6), and ET/base/CommonOperationsTest.java:21-24: applying
    # (0 4 1) to (0 4 5 6 1 2 3) gives (0 1 2 3 4 5 6).  So omega = (0 4 1)
    # with beta = (0 4 5 6 1 2 3) is a pair with target (0 ... 6) that the
    # single move (0 4 1) sorts.
    omega = [4, 0, 2, 3, 1, 5, 6]
    beta = [0, 4, 5, 6, 1, 2, 3]
    assert O.validate(beta, omega, [(0, 4, 1)], 1) == 0
    assert O.validate(beta, omega, [(0, 1, 4)], 1) != 0


def test_reference_codec_fixtures(O):
    # CT/base/CertificateCodecTest.java:14-38 on the 3-cycle pair
    omega, beta = [1, 2, 0], [0, 1, 2]
    key = O.codec_key(omega)
    assert len(key) == 2 and O.codec_decode_key(key, 3) == omega
    val = O.codec_value([(0, 1, 2)])
    assert len(val) == 2 and O.codec_decode_value(val) == [(0, 1, 2)]
    assert O.validate(beta, omega, [(0, 1, 2)], 1) == 0
    assert O.validate(beta, omega, [(0, 2, 1)], 1) != 0   # inapplicable
    assert O.validate(beta, omega, [(0, 1, 2)], 0) != 0   # overlong
    with pytest.raises(ValueError):
        O.codec_decode_value(b"\x00")
    with pytest.raises(ValueError):                      # nonzero padding
        O.codec_decode_value(bytes([val[0], val[1] | 1]))


def test_elias_hartman_reverse_distance(O):
    # d(reverse of 0..n-1) = floor(n/2) + 1 for n >= 3 (Elias and Hartman)
    for n in range(3, 11):
        beta, omega = O.linear_to_pair(O.pack(list(range(n - 1, -1, -1))), n)
        assert O.iddfs(beta, omega, O.MOVES_REF, 9)[0] == n // 2 + 1


def test_known_answers_reproduce(O):
    with open(os.path.join(GOLDEN, "expected.json")) as f:
        exp = json.load(f)
    for k in exp["known_answers"]:
        r = O.search(k["beta"], k["omega"], k["initial_even"], k["min_rate"], k["max_moves"],
                     True, O.MOVES_REF, O.FIRST)
        assert r.found == k["found"], k["src"]
        if r.found:
            sym = O.positions_to_symbols(k["beta"], r.witness())
            assert O.validate(k["beta"], k["omega"], sym, k["max_moves"]) == 0
            if "witness_symbols" in k:
                assert [list(s) for s in sym] == k["witness_symbols"]


def test_generators_match_manifest(O):
    import hashlib
    with open(os.path.join(GOLDEN, "instances.json")) as f:
        inst = json.load(f)

    def sha(*arrays):
        h = hashlib.sha256()
        for a in arrays:
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()

    c1 = O.gen_uniform(0, 10, 10000)
    assert sha(c1) == inst["config1"]["sha256"]
    # Durstenfeld shuffle outputs are permutations, never the identity
    for x in c1[:200]:
        p = O.unpack(x, 10)
        assert sorted(p) == list(range(10)) and p != list(range(10))
    ns, ks, p = O.gen_products(1, 12, 16, 4, 9, 100000)
    assert sha(ns, ks, p) == inst["config2"]["sha256"]
    assert ns.min() == 12 and ns.max() == 16 and ks.min() == 4 and ks.max() == 9
    # SplitMix64 reference stream (seed 0): the first output of the published
    # algorithm is 0xe220a8397b1dcdaf
    assert O.splitmix64_stream(0, 1)[0] == 0xE220A8397B1DCDAF


def greedy_search_py(beta, omega, max_depth, sat, reach):
    """Pure-Python restatement of the greedy FIRST order for small inputs:
    the reference filter (no fixed point, not three distinct cycles, delta 0
    or +2) with deltas by full recount, children of a node evaluated in
    lexicographic order (a satisfying one ends the search), expandable
    children visited gaining ones first, each class in lexicographic order."""
    n = len(beta)
    state = {"expanded": 0, "witness": None}

    def dfs(beta, omega, odd, depth, path):
        state["expanded"] += 1
        kids = []
        for i, j, k in itertools.combinations(range(n), 3):
            a, b, c = beta[i], beta[j], beta[k]
            if omega[a] == a or omega[b] == b or omega[c] == c:
                continue
            lab = cycle_sets(omega)
            if len({lab[a], lab[b], lab[c]}) == 3:
                continue
            child = recount_after(omega, a, b, c)
            codd = odd_cycles_py(child)
            if codd - odd not in (0, 2):
                continue
            if (sat[depth + 1] >> codd) & 1:
                state["witness"] = path + [(i, j, k)]
                return True
            if depth + 1 < max_depth and (reach[depth + 1] >> codd) & 1:
                kids.append((codd > odd, i, j, k, child, codd))
        for gain in (True, False):
            for g, i, j, k, child, codd in kids:
                if g != gain:
                    continue
                cb = beta[:i] + beta[j:k] + beta[i:j] + beta[k:]
                if dfs(cb, child, codd, depth + 1, path + [(i, j, k)]):
                    return True
        return False

    found = dfs(list(beta), list(omega), odd_cycles_py(omega), 0, [])
    return found, state["witness"], state["expanded"]


def test_greedy_order_matches_python_restatement(O):
    rng = random.Random(23)
    checked = 0
    for trial in range(120):
        n = rng.randint(4, 7)
        beta = list(range(n))
        rng.shuffle(beta)
        omega = list(range(n))
        rng.shuffle(omega)
        for k in (2, 3, 4):
            odd = O.count_odd_cycles(omega)
            sat, reach = O.rate_tables(n, odd, 0.0, k, True)
            r = O.search(beta, omega, odd, 0.0, k, True, O.MOVES_REF, O.FIRST, greedy=True)
            found, wit, expanded = greedy_search_py(beta, omega, k, sat, reach)
            assert bool(r.found) == found
            assert r.expanded == expanded
            if found:
                assert r.witness() == wit
                checked += 1
    assert checked > 20


def test_greedy_order_is_order_independent_where_it_must_be(O):
    # exhaustive searches (no witness, or COUNT_ALL) expand the same node set
    # in either order; FIRST found / not-found never depends on the order
    ns, ks, perms = O.gen_products(5, 8, 10, 3, 5, 80)
    nonexist = 0
    for q in range(len(ns)):
        beta, omega = O.linear_to_pair(int(perms[q]), int(ns[q]))
        odd = O.count_odd_cycles(omega)
        k = int(ks[q])
        for bound in (k, max(1, k - 2)):
            lex = O.search(beta, omega, odd, 0.0, bound, True)
            grd = O.search(beta, omega, odd, 0.0, bound, True, greedy=True)
            assert lex.found == grd.found
            if not lex.found:
                nonexist += 1
                assert lex.expanded == grd.expanded
            else:
                sym = O.positions_to_symbols(beta, grd.witness())
                assert O.validate(beta, omega, sym, bound) == 0
        ca = O.search(beta, omega, odd, 0.0, k, True, O.MOVES_REF, O.COUNT_ALL)
        cg = O.search(beta, omega, odd, 0.0, k, True, O.MOVES_REF, O.COUNT_ALL, greedy=True)
        assert (ca.expanded, ca.solutions) == (cg.expanded, cg.solutions)
    assert nonexist > 10


def test_full_size_fixtures_are_the_oracles(O):
    """The committed full-size answers (tests/golden/make_golden.py) are the
    oracle's: all 10,000 n = 10 distances recomputed in both move sets, with
    the odd-cycle bound <= d <= d(10) = 6, the reference filter never beating
    the full move set and the distance histogram of the committed set; the
    n = 16 depth-8 found / not-found on every 25th instance, the four
    non-existence proofs at the positions the proofs fixture names, and the
    exhaustive counts' tree sizes under the recorded cap."""
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                           "expected.json")) as f:
        exp = json.load(f)
    c1 = O.gen_uniform(0, 10, 10000)
    ten = np.full(10000, 10, dtype=np.uint8)
    nine = np.full(10000, 9, dtype=np.uint8)
    dist = {}
    for key, ms in (("all", O.MOVES_ALL), ("ref", O.MOVES_REF)):
        res = O.search_batch_ex(ten, c1, nine, ms, O.FIRST, threads=8, iddfs=True)
        d = np.array([r.depth for r in res])
        assert "".join(map(str, d)) == exp["config1_distances"][key]
        dist[key] = d
    assert (dist["all"] <= dist["ref"]).all()
    for q in range(0, 10000, 97):
        beta, omega = O.linear_to_pair(c1[q], 10)
        lb = (11 - O.count_odd_cycles(omega)) // 2
        assert lb <= dist["all"][q] <= 6
    assert list(np.bincount(dist["all"], minlength=7)) == [0, 0, 30, 477, 3268, 5536, 689]
    # n = 16, move bound 8
    _, _, p3 = O.gen_products(2, 16, 16, 9, 9, 1999)
    perms = np.concatenate([p3, np.array([O.pack(list(range(15, -1, -1)))], dtype=np.uint64)])
    idx = np.arange(0, 2000, 25)
    res = O.search_batch_ex(np.full(len(idx), 16, np.uint8), perms[idx],
                            np.full(len(idx), 8, np.uint8), O.MOVES_REF, O.FIRST, threads=8,
                            greedy=True)
    assert [r.found for r in res] == [int(exp["config3_found"][q]) for q in idx]
    assert [q for q, c in enumerate(exp["config3_found"]) if c == "0"] == \
        [p["q"] for p in exp["config3_nonexistence"]]
    ca = exp["config3_count_all"]
    assert all(r["expanded"] <= ca["cap"] for r in ca["rows"]) and len(ca["rows"]) <= ca["zero_slack"]
    for r in ca["rows"]:
        _, omega = O.linear_to_pair(r["perm"], 16)
        assert (17 - O.count_odd_cycles(omega)) // 2 == 8  # zero slack at bound 8
