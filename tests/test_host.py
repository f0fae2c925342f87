"""CPU tests of libtd16.so's host side and of the Python mirrors of the
reference's dispatcher classes.  No GPU call is made here."""
import json
import os
import random
import re

import numpy as np
import pytest

from td16_amd import Td16Error
from td16_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_library_exports_every_declared_symbol(native):
    with open(os.path.join(ROOT, "include", "td16_abi.h")) as f:
        header = f.read()
    declared = set(re.findall(r"\b(td16_[a-z0-9_]+)\s*\(", header))
    assert len(declared) >= 25
    lib = _lib.lib()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    # and the binding covers all of them
    assert declared == set(_lib.SIGNATURES)


def test_no_fallback_when_library_missing(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(ImportError):
        _lib.lib()


def test_generators_match_oracle(O, native):
    for seed, n in ((0, 10), (4, 7), (9, 16)):
        assert np.array_equal(native.gen_uniform(seed, n, 500), O.gen_uniform(seed, n, 500))
    a = native.gen_products(1, 12, 16, 4, 9, 3000)
    b = O.gen_products(1, 12, 16, 4, 9, 3000)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    with open(os.path.join(GOLDEN, "instances.json")) as f:
        inst = json.load(f)
    assert [int(x) for x in native.gen_uniform(0, 10, 8)] == inst["config1"]["first"]
    gaps = native.gen_exponential(3, 20000)
    assert abs(gaps.mean() - 1.0) < 0.03 and gaps.min() >= 0.0


def test_linear_request_matches_oracle_pair(O, native):
    ns, ks, perms = O.gen_products(8, 3, 16, 1, 9, 200)
    for q in range(200):
        req = native.linear_request(int(perms[q]), int(ns[q]), int(ks[q]))
        beta, omega = O.linear_to_pair(perms[q], int(ns[q]))
        assert list(req.beta[:req.n]) == beta and list(req.omega[:req.n]) == omega
        assert req.initial_even == O.count_odd_cycles(omega)
        t = (req.n - req.initial_even) // 2
        assert req.min_rate == float(np.float32(2) / (np.float32(ks[q]) / np.float32(t)))


def test_validator_and_positions_match_oracle(O, native):
    rng = random.Random(5)
    ns, ks, perms = O.gen_products(12, 6, 12, 2, 6, 150)
    for q in range(150):
        beta, omega = O.linear_to_pair(perms[q], int(ns[q]))
        r = O.search(beta, omega, O.count_odd_cycles(omega), 0.0, int(ks[q]), True)
        assert r.found == 1
        pos = r.witness()
        sym = native.positions_to_symbols(beta, pos)
        assert sym == O.positions_to_symbols(beta, pos)
        native.validate(beta, omega, sym, int(ks[q]))            # valid: no raise
        assert O.validate(beta, omega, sym, int(ks[q])) == 0
        # corruptions the reference rejects: too long, permuted triple, truncated
        with pytest.raises(Td16Error):
            native.validate(beta, omega, sym, len(sym) - 1)
        a, b, c = sym[0]
        bad = [(a, c, b)] + list(sym[1:])
        with pytest.raises(Td16Error):
            native.validate(beta, omega, bad, int(ks[q]))
        assert O.validate(beta, omega, bad, int(ks[q])) != 0
        if len(sym) > 1:
            with pytest.raises(Td16Error):
                native.validate(beta, omega, sym[:-1], int(ks[q]))
            assert O.validate(beta, omega, sym[:-1], int(ks[q])) != 0
    with pytest.raises(Td16Error):
        native.validate([0, 1, 2], [1, 2, 0], [], 1)


def test_codec_matches_oracle(O, native):
    rng = random.Random(2)
    for _ in range(300):
        n = rng.randint(1, 17)
        om = list(range(n))
        rng.shuffle(om)
        assert native.codec_key(om) == O.codec_key(om)
        assert native.codec_decode_key(native.codec_key(om), n) == om
        m = rng.randint(1, 9)
        moves = [tuple(rng.sample(range(17), 3)) for _ in range(m)]
        v = native.codec_value(moves)
        assert v == O.codec_value(moves) and len(v) == (15 * m + 7) // 8
        assert native.codec_decode_value(v) == moves
    # CT/base/CertificateCodecTest.java:14-38
    assert len(native.codec_key([1, 2, 0])) == 2
    assert len(native.codec_value([(0, 1, 2)])) == 2
    with pytest.raises(Td16Error):
        native.codec_decode_value(b"\x00")
    with pytest.raises(Td16Error):
        native.codec_value([])
    with pytest.raises(Td16Error):
        native.codec_value([(0, 1, 32)])


def test_codec_bytes_written_out_by_hand(native):
    """The certificate layout (CertificateCodec.java:22-70): each symbol in 5
    bits, most significant bit first, concatenated and zero-padded to a
    byte.  These bytes were worked out on paper from that definition, so a
    mistake shared by the product and the oracle (bit order, padding side)
    fails here."""
    # [1, 2, 0] -> 00001 00010 00000 + 0 -> 0000 1000 | 1000 0000
    assert native.codec_key([1, 2, 0]) == bytes([0x08, 0x80])
    # (0, 1, 2) -> 00000 00001 00010 + 0 -> 0000 0000 | 0100 0100
    assert native.codec_value([(0, 1, 2)]) == bytes([0x00, 0x44])
    # 17 symbols, the reverse 16..0: 85 bits and 3 padding bits in 11 bytes
    rev = list(range(16, -1, -1))
    key = bytes.fromhex("83 dc d6 2d 49 41 cc 52 0c 41 00")
    assert native.codec_key(rev) == key
    assert native.codec_decode_key(key, 17) == rev
    # two moves: 10000 00000 00101 00011 01001 01100 + 00
    val = bytes.fromhex("80 0a 34 b0")
    assert native.codec_value([(16, 0, 5), (3, 9, 12)]) == val
    assert native.codec_decode_value(val) == [(16, 0, 5), (3, 9, 12)]
    # a set padding bit is rejected by both decoders
    with pytest.raises(Td16Error):
        native.codec_decode_key(bytes([0x08, 0x81]), 3)
    with pytest.raises(Td16Error):
        native.codec_decode_value(bytes([0x00, 0x45]))


# oracle replay codes -> td16_check_batch / td16_certify_batch status
_ORACLE_TO_STATUS = {0: 0, -1: -1, -6: -2, -2: -3, -3: -4, -4: -5, -5: -6}


def _oracle_results(O, native, seed, count):
    ns, ks, perms = O.gen_products(seed, 5, 12, 1, 6, count)
    reqs, results, syms = [], [], []
    for q in range(count):
        req = native.linear_request(int(perms[q]), int(ns[q]), int(ks[q]))
        beta, omega = list(req.beta[:req.n]), list(req.omega[:req.n])
        # every third request gets a tighter bound, so some have no witness
        mm = int(ks[q]) - (1 if q % 3 == 0 and ks[q] > 1 else 0)
        req.max_moves = mm
        r = O.search(beta, omega, req.initial_even, req.min_rate, mm, True)
        res = _lib.Result()
        res.found = r.found
        w = r.witness() if r.found else []
        res.depth = len(w)
        for d, mv in enumerate(w):
            res.moves[d][0], res.moves[d][1], res.moves[d][2] = mv
        reqs.append(req)
        results.append(res)
        syms.append(O.positions_to_symbols(beta, w) if w else [])
    return reqs, results, syms


def test_certify_batch_matches_oracle_codec_and_replay(O, native):
    reqs, results, syms = _oracle_results(O, native, 31, 240)
    # a device error and an invalid position triple are not certified
    results[5].status = _lib.DEVICE_ERROR
    bad = next(q for q in range(6, 240) if results[q].found and results[q].depth >= 1)
    results[bad].moves[0][0], results[bad].moves[0][1] = results[bad].moves[0][1], \
        results[bad].moves[0][0]
    for threads in (1, 3, 0):
        certs = native.certify_batch(reqs, results, threads)
        for q, req in enumerate(reqs):
            n = req.n
            beta, omega = list(req.beta[:n]), list(req.omega[:n])
            if q == 5 or not results[q].found:
                assert certs.status[q] == 1 and certs.value_len[q] == 0
                continue
            if q == bad:
                assert certs.status[q] == -4 and certs.value_len[q] == 0
                continue
            assert certs.status[q] == 0
            assert O.validate(beta, omega, syms[q], req.max_moves) == 0
            assert certs.key(q, n) == O.codec_key(omega)
            assert certs.value(q) == O.codec_value(syms[q])
            assert O.codec_decode_value(certs.value(q)) == syms[q]
    assert (certs.status == 1).sum() >= 10 and (certs.status == 0).sum() >= 100
    empty = native.certify_batch([], [])
    assert len(empty.status) == 0


def test_check_batch_matches_oracle_replay(O, native):
    reqs, results, syms = _oracle_results(O, native, 32, 300)
    certs = native.certify_batch(reqs, results)
    count = len(reqs)
    betas = np.zeros((count, _lib.MAX_N), np.uint8)
    omegas = np.zeros((count, _lib.MAX_N), np.uint8)
    ns = np.array([r.n for r in reqs], np.int32)
    mm = np.array([r.max_moves for r in reqs], np.int32)
    for q, r in enumerate(reqs):
        betas[q, :r.n] = r.beta[:r.n]
        omegas[q, :r.n] = r.omega[:r.n]
    values = certs.values.copy()
    vlen = certs.value_len.copy()
    rng = random.Random(4)
    expect_malformed = set()
    stored = [q for q in range(count) if vlen[q] > 0]
    for q in stored:
        kind = q % 7
        moves = O.codec_decode_value(bytes(values[q, :vlen[q]]))
        if kind == 1:                      # a permuted triple
            a, b, c = moves[0]
            moves[0] = (a, c, b)
        elif kind == 2 and len(moves) > 1:  # truncated certificate
            moves = moves[:-1]
        elif kind == 3:                    # a foreign symbol
            moves[-1] = (moves[-1][0], moves[-1][1], rng.randint(ns[q], 31))
        elif kind == 4:                    # the bound tightened below the length
            mm[q] = len(moves) - 1
        elif kind == 5:                    # nonzero padding bit
            nb = int(vlen[q])
            pad = 8 * nb - 15 * len(moves)
            if pad:
                values[q, nb - 1] |= 1
                expect_malformed.add(q)
            continue
        elif kind == 6:                    # length that is no whole number of moves
            vlen[q] = 1
            expect_malformed.add(q)
            continue
        v = O.codec_value(moves)
        values[q, :] = 0
        values[q, :len(v)] = list(v)
        vlen[q] = len(v)
    st = native.check_batch(betas, omegas, ns, mm, values, vlen, threads=4)
    seen = set()
    for q in range(count):
        if vlen[q] == 0:
            assert st[q] == 1
        elif q in expect_malformed:
            assert st[q] == -7
            with pytest.raises(ValueError):
                O.codec_decode_value(bytes(values[q, :vlen[q]]))
        else:
            n = int(ns[q])
            moves = O.codec_decode_value(bytes(values[q, :vlen[q]]))
            code = O.validate(list(betas[q, :n]), list(omegas[q, :n]), moves, int(mm[q]))
            assert st[q] == _ORACLE_TO_STATUS[code], (q, code, st[q])
        seen.add(int(st[q]))
    assert {0, 1, -1, -3, -4, -7} <= seen
    assert len(native.check_batch(betas[:0], omegas[:0], ns[:0], 9, values[:0], vlen[:0])) == 0


def test_emitter_stream_generator_and_obligation_view(O, native):
    def unpack(v, n):
        return [(int(v) >> (4 * q)) & 0xF for q in range(n)]

    rev = native.gen_reverse_like(3, 16, 300, rmax=0)
    assert all(unpack(v, 16) == list(range(15, -1, -1)) for v in rev)
    short = native.gen_reverse_like(3, 12, 5, rmax=0)
    assert all(int(v) >> 48 == 0xFFFF for v in short)           # unused nibbles stay 0xF
    ns, ks, perms = native.gen_emitter_stream(5, 2000)
    assert set(ns) == {16} and set(ks) == {9}
    for q in range(2000):
        p = unpack(perms[q], 16)
        assert sorted(p) == list(range(16))
        if q % 10 == 9:
            # reverse-like: at most two block transpositions from the reverse,
            # so at least 16 - 2*3 adjacent pairs stay descending
            assert sum(p[i] == p[i + 1] + 1 for i in range(15)) >= 15 - 6
    _, _, easy = native.gen_products(5, 16, 16, 9, 9, 1800)
    assert np.array_equal(perms[np.arange(2000) % 10 != 9], easy)
    reqs = [native.linear_request(int(perms[q]), 16, 9) for q in range(50)]
    arr = (_lib.Request * 50)(*reqs)
    betas, omegas, nn, mm = native.obligations(arr)
    for q, r in enumerate(reqs):
        assert nn[q] == r.n == 17 and mm[q] == 9
        assert list(betas[q, :17]) == list(r.beta) and list(omegas[q, :17]) == list(r.omega)
        beta, omega = O.linear_to_pair(perms[q], 16)
        assert list(betas[q, :17]) == beta and list(omegas[q, :17]) == omega


def test_search_argument_errors_without_gpu(native):
    # jni_sorter.cu:61-67, sorting_search.cu:820-825: argument errors surface
    # before any device work
    with pytest.raises(Td16Error) as e:
        native.search_positions(0, list(range(18)), list(range(18)), 1, 1.0, 3)
    assert e.value.status == _lib.INVALID_ARGUMENT
    with pytest.raises(Td16Error) as e:
        native.search_positions(0, [0, 1, 2], [1, 2], 1, 1.0, 1)
    assert e.value.status == _lib.INVALID_ARGUMENT
    with pytest.raises(Td16Error) as e:
        native.init_slots(2, [], 1 << 20)
    assert e.value.status == _lib.INVALID_ARGUMENT
    with pytest.raises(Td16Error) as e:
        native.init_slots(0, [0], 1 << 20)
    assert e.value.status == _lib.INVALID_ARGUMENT
    with pytest.raises(Td16Error) as e:
        native.init_slots(1, [0], 0)
    assert e.value.status == _lib.INVALID_ARGUMENT
    req = native.make_request([0, 1, 2], [1, 2, 0], 1, 2.0, 1)
    with pytest.raises(Td16Error) as e:
        native.search_ex(0, req)
    assert e.value.status == _lib.NOT_INITIALIZED


def test_dispatcher_rejects_bad_device_lists_before_any_gpu_call():
    from td16_amd import GPUSortingSearch
    for devices, slots in (([0, 0], 7), ([-1], 7), ([], 7), ([0], 0)):
        with pytest.raises(ValueError):
            GPUSortingSearch(devices, slots, 1 << 20)
    with pytest.raises(ValueError):
        GPUSortingSearch([0], 2, 0)


def test_autotune_rejects_bad_arguments_before_any_launch(native):
    req = native.make_request([0, 1, 2], [1, 2, 0], 1, 2.0, 1, True, native.MOVES_REF,
                              native.FIRST)
    with pytest.raises(Td16Error, match="non-empty batch"):
        native.autotune([])
    with pytest.raises(Td16Error, match="repeats"):
        native.autotune([req], repeats=0)
    with pytest.raises(Td16Error, match="repeats"):
        native.autotune([req], repeats=101)


def test_bench_strided_sample_copies_every_kth_request(native):
    import bench
    ns, ks, perms = native.gen_products(1, 12, 16, 4, 9, 23)
    reqs = bench.build_requests(native, ns, ks, perms)
    sample = bench.strided_sample(reqs, 5)
    assert len(sample) == 5
    for i, q in enumerate(range(0, 23, 5)):
        assert bytes(sample[i]) == bytes(reqs[q])


def test_noalloc_rejects_oversized_batches_before_touching_a_gpu(native):
    # the size and option checks come before any HIP call, so they hold here
    reqs = (_lib.Request * (_lib.NOALLOC_MAX_REQUESTS + 1))()
    with pytest.raises(Td16Error) as e:
        native.batch_noalloc(reqs)
    assert e.value.status == _lib.INVALID_ARGUMENT
    assert "32768" in str(e.value) and "maximum" in str(e.value)
    with pytest.raises(Td16Error) as e:
        native.batch_noalloc(reqs[:4], use_graph=True)
    assert e.value.status == _lib.INVALID_ARGUMENT and "use_graph" in str(e.value)
