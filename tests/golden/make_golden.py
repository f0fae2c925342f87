"""Writes the golden fixtures in this folder from the CPU oracle.

    python tests/golden/make_golden.py

Outputs (all small JSON):
  instances.json  -- SHA-256 of every benchmark instance set (SplitMix64
                     generators, nibble-packed uint64 little-endian, plus the
                     n and k bytes where they vary), and the first instances
                     of each set in clear
  expected.json   -- oracle results: the exact distance of all 10,000 n = 10
                     permutations in both move sets, first witnesses and node
                     counts (mixes), found / not-found of all 2,000 n = 16
                     instances at depth 8, the depth-8 non-existence proofs
                     with their exact node counts, exhaustive depth-8 counts
                     of the zero-slack n = 16 instances whose trees the CPU
                     finishes, COUNT_ALL counts of small products, and the
                     reference's own known-answer pairs.
The reference code is never imported or run: its fixtures enter as literals
copied from its test files and logs (cited next to each).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

THREADS = max(1, min(16, os.cpu_count() or 1))
# exhaustive trees larger than this are left to size-independent checks
ZERO_SLACK_CAP = 60_000_000


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def cycles_to_oneline(n, cycles):
    om = list(range(n))
    for cyc in cycles:
        for i, x in enumerate(cyc):
            om[x] = cyc[(i + 1) % len(cyc)]
    return om


def witness_rows(r):
    return [list(m) for m in r.witness()]


def result_dict(r, with_witness=True):
    d = {"found": int(r.found), "depth": int(r.depth), "expanded": int(r.expanded),
         "solutions": int(r.solutions)}
    if with_witness:
        d["witness"] = witness_rows(r)
    return d


def linear_case(perm, n, k, moveset=O.MOVES_REF, mode=O.FIRST):
    beta, omega = O.linear_to_pair(perm, n)
    N = n + 1
    odd = O.count_odd_cycles(omega)
    target = (N - odd) // 2
    rate = float(np.float32(2.0) / (np.float32(k) / np.float32(target)))
    return beta, omega, odd, rate


def main():
    inst = {}
    # config 1: 10,000 uniform permutations of n = 10, seed 0
    c1 = O.gen_uniform(0, 10, 10000)
    inst["config1"] = {"seed": 0, "n": 10, "count": 10000, "sha256": sha(c1),
                       "first": [int(x) for x in c1[:8]]}
    # config 2: 100,000 products, n U{12..16}, k U{4..9}, seed 1
    ns2, ks2, p2 = O.gen_products(1, 12, 16, 4, 9, 100000)
    inst["config2"] = {"seed": 1, "count": 100000, "sha256": sha(ns2, ks2, p2),
                       "first": [[int(ns2[q]), int(ks2[q]), int(p2[q])] for q in range(8)]}
    # config 3: 1,999 products of 9 at n = 16 (seed 2) plus the reverse permutation
    ns3, ks3, p3 = O.gen_products(2, 16, 16, 9, 9, 1999)
    rev = O.pack(list(range(15, -1, -1)))
    p3full = np.concatenate([p3, np.array([rev], dtype=np.uint64)])
    inst["config3"] = {"seed": 2, "count": 2000, "sha256": sha(p3full),
                       "first": [int(x) for x in p3full[:8]], "reverse": int(rev)}
    # config 4: served stream instances (seed 3, k U{6..9}) and unit-rate gaps
    ns4, ks4, p4 = O.gen_products(3, 12, 16, 6, 9, 20000)
    inst["config4"] = {"seed": 3, "count": 20000, "sha256": sha(ns4, ks4, p4)}
    with open(os.path.join(HERE, "instances.json"), "w") as f:
        json.dump(inst, f, indent=1)

    exp = {}
    # config 1: exact distances of the first 300 by IDDFS over ALL and REF moves
    rows = []
    for q in range(300):
        beta, omega = O.linear_to_pair(c1[q], 10)
        d_all, _ = O.iddfs(beta, omega, O.MOVES_ALL, 9)
        d_ref, _ = O.iddfs(beta, omega, O.MOVES_REF, 9)
        rows.append([int(c1[q]), d_all, d_ref])
    exp["config1_iddfs"] = rows
    # config 1 in full: one digit per permutation, both move sets
    ten = np.full(10000, 10, dtype=np.uint8)
    nine = np.full(10000, 9, dtype=np.uint8)
    for name, ms in (("all", O.MOVES_ALL), ("ref", O.MOVES_REF)):
        res = O.search_batch_ex(ten, c1, nine, ms, O.FIRST, threads=THREADS, iddfs=True)
        assert all(r.found == 1 for r in res)
        exp.setdefault("config1_distances", {})[name] = "".join(str(r.depth) for r in res)
    # config 2: first 600 FIRST-mode REF searches
    rows = []
    for q in range(600):
        beta, omega, odd, rate = linear_case(p2[q], int(ns2[q]), int(ks2[q]))
        r = O.search(beta, omega, odd, rate, int(ks2[q]), True, O.MOVES_REF, O.FIRST)
        rows.append({"q": q, **result_dict(r)})
    exp["config2_first"] = rows
    # config 3: depth-8 FIRST searches of the first 150, and the four
    # non-existence proofs (complete trees: node counts are exact)
    rows = []
    for q in list(range(150)):
        beta, omega, odd, rate = linear_case(p3full[q], 16, 8)
        r = O.search(beta, omega, odd, rate, 8, True, O.MOVES_REF, O.FIRST)
        rows.append({"q": q, **result_dict(r)})
    exp["config3_first"] = rows
    proofs = []
    for q in (220, 594, 1864, 1999):
        beta, omega, odd, rate = linear_case(p3full[q], 16, 8)
        r = O.search(beta, omega, odd, rate, 8, True, O.MOVES_REF, O.FIRST)
        assert r.found == 0
        proofs.append({"q": q, "perm": int(p3full[q]), **result_dict(r, False)})
    exp["config3_nonexistence"] = proofs
    # config 3 in full: found (1) / not found (0) within 8 moves, REF, FIRST
    sixteen = np.full(2000, 16, dtype=np.uint8)
    eight = np.full(2000, 8, dtype=np.uint8)
    res = O.search_batch_ex(sixteen, p3full, eight, O.MOVES_REF, O.FIRST, threads=THREADS,
                            greedy=True)
    assert all(r.found in (0, 1) for r in res)
    exp["config3_found"] = "".join(str(r.found) for r in res)
    assert [q for q, r in enumerate(res) if r.found == 0] == [p["q"] for p in proofs]
    # exhaustive depth-8 counts (COUNT_ALL) of the zero-slack instances: the
    # odd-cycle bound of the root is 8, so only moves that gain two odd
    # cycles stay reachable and the tree is the same under ALL and REF moves;
    # the instances whose trees stay within ZERO_SLACK_CAP nodes
    zero = [q for q in range(2000)
            if (17 - O.count_odd_cycles(O.linear_to_pair(p3full[q], 16)[1])) // 2 == 8]
    res = O.search_batch_ex(sixteen[:len(zero)], p3full[zero], eight[:len(zero)], O.MOVES_REF,
                            O.COUNT_ALL, threads=THREADS, cap=ZERO_SLACK_CAP)
    exp["config3_count_all"] = {
        "cap": ZERO_SLACK_CAP, "zero_slack": len(zero),
        "rows": [{"q": q, "perm": int(p3full[q]), "expanded": int(r.expanded),
                  "solutions": int(r.solutions)}
                 for q, r in zip(zero, res) if r.found != -2]}
    # COUNT_ALL: zero- and one-slack bounds on small products, both move sets
    rows = []
    ns5, ks5, p5 = O.gen_products(5, 8, 11, 4, 6, 40)
    for q in range(40):
        for ms in (O.MOVES_REF, O.MOVES_ALL):
            beta, omega = O.linear_to_pair(p5[q], int(ns5[q]))
            odd = O.count_odd_cycles(omega)
            lb = (len(beta) - odd) // 2
            for slack in ((0, 1) if int(ns5[q]) <= 8 or ms == O.MOVES_REF else (0,)):
                D = min(9, lb + slack)
                rate = float(np.float32(2.0) / (np.float32(D) / np.float32(lb)))
                r = O.search(beta, omega, odd, rate, D, True, ms, O.COUNT_ALL)
                if r.expanded > 300000:
                    continue
                rows.append({"perm": int(p5[q]), "n": int(ns5[q]), "moveset": ms, "D": D,
                             "rate": rate, "odd": odd, **result_dict(r, False)})
    exp["count_all"] = rows

    # known-answer pairs from the reference's own tests and logs
    kap = []
    # GPUSortingSearchSmokeTest.java:26-35
    kap.append({"src": "ET/base/GPUSortingSearchSmokeTest.java:26-35", "beta": [0, 1, 2],
                "omega": [1, 2, 0], "initial_even": 1, "min_rate": 2.0, "max_moves": 1,
                "found": 1, "witness_symbols": [[0, 1, 2]]})
    # CT/OddLengthOrientedCyclesTest.java:84-86; searched with maxMoves 7,
    # minRate 2/(7/6f) (WitnessEmitter.java:72, OddLengthOrientedCycles.java:66-67);
    # all satisfied (emitter.log:74)
    for cyc in ([0, 7, 1, 8, 2, 9, 3, 10, 4, 11, 5, 12, 6],
                [0, 10, 7, 4, 1, 11, 8, 5, 2, 12, 9, 6, 3],
                [0, 11, 9, 7, 5, 3, 1, 12, 10, 8, 6, 4, 2]):
        om = cycles_to_oneline(13, [cyc])
        rate = float(np.float32(2.0) / (np.float32(7.0) / np.float32(6.0)))
        kap.append({"src": "CT/OddLengthOrientedCyclesTest.java:84-86; emitter.log:74",
                    "beta": list(range(13)), "omega": om, "initial_even": O.count_odd_cycles(om),
                    "min_rate": rate, "max_moves": 7, "found": 1})
    # emitter.log:71-72: the oriented 9-cycle left unsatisfied at maxMoves 5,
    # minRate 2/(5/4f) (WitnessEmitter.java:70)
    om = cycles_to_oneline(9, [[0, 5, 1, 6, 2, 7, 3, 8, 4]])
    rate = float(np.float32(2.0) / (np.float32(5.0) / np.float32(4.0)))
    kap.append({"src": "emitter.log:71-72", "beta": list(range(9)), "omega": om,
                "initial_even": O.count_odd_cycles(om), "min_rate": rate, "max_moves": 5,
                "found": 0})
    # ET/DirectTwistedBraceletGeneratorTest.java:168: a [7,4,2,2,2] representative,
    # requiredIncrease 8, maxMoves 9 (emitter.log:29-30: all 4,800 classes solved)
    om = cycles_to_oneline(17, [[0, 13, 11, 8, 6, 4, 2], [1, 10, 5, 3], [7, 15], [9, 16],
                                [12, 14]])
    rate = float(np.float32(2.0) / (np.float32(9.0) / np.float32(8.0)))
    kap.append({"src": "ET/DirectTwistedBraceletGeneratorTest.java:168; emitter.log:30",
                "beta": list(range(17)), "omega": om, "initial_even": O.count_odd_cycles(om),
                "min_rate": rate, "max_moves": 9, "found": 1})
    for k in kap:
        r = O.search(k["beta"], k["omega"], k["initial_even"], k["min_rate"], k["max_moves"],
                     True, O.MOVES_REF, O.FIRST)
        assert r.found == k["found"], k["src"]
        k["oracle"] = result_dict(r)
    exp["known_answers"] = kap
    with open(os.path.join(HERE, "expected.json"), "w") as f:
        json.dump(exp, f)
    print("wrote", os.path.join(HERE, "instances.json"), os.path.join(HERE, "expected.json"))


if __name__ == "__main__":
    main()
