/* Runs the CPU oracle (oracle/td16_oracle.c) under sanitizers:
 * tests/test_native_host.py builds it with gcc under AddressSanitizer +
 * UndefinedBehaviorSanitizer and under ThreadSanitizer (the batch runner is
 * threaded).  Checks a few facts the oracle's own tests pin in full
 * (BFS diameter, IDDFS = BFS, replay of every witness), so that the runs
 * exercise every code path.  Prints "oracle selftest ok". */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/td16_oracle.h"

static int fails = 0;
#define CHECK(c)                                                      \
    do {                                                              \
        if (!(c)) {                                                   \
            fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++fails;                                                  \
        }                                                             \
    } while (0)

int main(void) {
    /* BFS over S_6 and iterative deepening agree on every permutation */
    uint8_t *dist = malloc(720);
    const int diam = td16o_bfs(6, dist, NULL);
    CHECK(diam == 4);
    for (uint64_t r = 0; r < 720; r += 7) {
        uint8_t p[16], beta[17], omega[17];
        td16o_perm_unrank(r, 6, p);
        uint64_t packed = ~0ULL;
        for (int q = 0; q < 6; ++q) {
            packed &= ~(0xFULL << (4 * q));
            packed |= (uint64_t)p[q] << (4 * q);
        }
        const int N = td16o_linear_to_pair(packed, 6, beta, omega);
        td16o_result res;
        const int d = td16o_iddfs(beta, omega, N, TD16O_MOVES_ALL, 9, &res);
        CHECK(d == (dist[r] ? dist[r] : -1) || (dist[r] == 0 && d == 0));
    }
    free(dist);
    /* searches in both orders and modes, witnesses replayed, codec round trip */
    uint8_t ns[200], ks[200];
    uint64_t perms[200];
    td16o_gen_products(5, 8, 16, 3, 8, 200, ns, ks, perms);
    for (int q = 0; q < 200; ++q) {
        uint8_t beta[17], omega[17];
        const int N = td16o_linear_to_pair(perms[q], ns[q], beta, omega);
        const int odd = td16o_count_odd_cycles(omega, N);
        const float rate = 2.0f / ((float)ks[q] / (float)((N - odd) / 2));
        td16o_result res;
        CHECK(td16o_search_order(beta, omega, N, odd, rate, ks[q], 1, TD16O_MOVES_REF,
                                 TD16O_FIRST, q & 1, &res) == 0);
        if (res.found) {
            int32_t sym[27];
            uint8_t buf[32];
            int32_t back[96];
            CHECK(td16o_positions_to_symbols(beta, N, &res.moves[0][0], res.depth, sym) == 0);
            CHECK(td16o_validate(beta, omega, N, sym, res.depth, ks[q]) == 0);
            const int nb = td16o_codec_value(sym, res.depth, buf, 32);
            CHECK(nb > 0 && td16o_codec_decode_value(buf, nb, back, 32) == res.depth);
            CHECK(memcmp(back, sym, sizeof(int32_t) * 3 * (size_t)res.depth) == 0);
        }
    }
    /* the threaded batch runner, capped, iterative deepening and exhaustive */
    td16o_result out[200];
    CHECK(td16o_search_batch_ex(200, ns, perms, ks, TD16O_MOVES_REF, TD16O_FIRST, 1, 0, 0, 8,
                                out) == 0);
    for (int q = 0; q < 200; ++q) CHECK(out[q].found == 1);
    CHECK(td16o_search_batch_ex(60, ns, perms, ks, TD16O_MOVES_ALL, TD16O_FIRST, 0, 1, 0, 8,
                                out) == 0);
    for (int q = 0; q < 60; ++q) CHECK(out[q].found == 1 && out[q].depth <= ks[q]);
    CHECK(td16o_search_batch_ex(40, ns, perms, ks, TD16O_MOVES_REF, TD16O_COUNT_ALL, 0, 0, 5000,
                                8, out) == 0);
    if (fails) {
        printf("oracle selftest: %d failures\n", fails);
        return 1;
    }
    printf("oracle selftest ok\n");
    return 0;
}
