/*
 * td16 CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's witness search (symbol form:
 * a cyclic sequence beta and a one-line permutation omega), its certificate
 * replay and its certificate codec.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker.  The
 * product (td16_amd/libtd16.so) never links, loads or calls it.
 *
 * Reference files (paths relative to the reference root):
 *   N/ = emitter/src/native/, E/ = emitter/src/main/java/br/unb/cic/tdp/
 *
 * Pinning of every function (how each is checked independently of the
 * restatement itself) is listed next to each declaration.
 */
#ifndef TD16_ORACLE_H
#define TD16_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TD16O_MAX_N 17
#define TD16O_MAX_DEPTH 9
#define TD16O_DELTA_SKIP (-1000)

enum { TD16O_MOVES_REF = 0, TD16O_MOVES_ALL = 1 };
enum { TD16O_FIRST = 0, TD16O_COUNT_ALL = 1 };

/* Number of odd-length cycles of omega, fixed points included.
 * Follows N/backtracking/sorting_search.cu:211-226.
 * Pinned: brute-force parity identity c_odd = n (mod 2); identity -> n;
 * n-cycle -> n&1 (tests/test_oracle.py). */
int td16o_count_odd_cycles(const uint8_t *omega, int n);

/* Change in odd-cycle count for omega' = omega o (a b c)^-1, or
 * TD16O_DELTA_SKIP when a, b, c lie on three distinct cycles.
 * Follows N/backtracking/sorting_search.cu:240-332.
 * Pinned: equals a full recount of the product for every triple of every
 * permutation with n <= 7 (tests/test_oracle.py). */
int td16o_delta(const uint8_t *omega, int a, int b, int c);

/* Witness and prune predicates as bit tables over the odd-cycle count:
 * bit e of sat[d] is satisfies_rate(e, d), bit e of reach[d] is
 * can_still_reach(e, d), for d = 0..9, in IEEE single precision exactly as
 * N/backtracking/sorting_search.cu:336-373 evaluates them.
 * Pinned: closed forms for the emitter's minRate convention
 * (E/UnorientedCycles.java:34) and hand-computed non-fullSorting cases. */
void td16o_rate_tables(int n, int initialEven, float minRate, int maxDepth,
                       int fullSorting, uint32_t sat[10], uint32_t reach[10]);

typedef struct {
    int32_t found;        /* 1 when a witness was found */
    int32_t depth;        /* witness length (0 when none) */
    int32_t moves[9][3];  /* positions (i, j, k) in the beta current before each move */
    uint64_t expanded;    /* nodes whose children were evaluated, root included */
    uint64_t solutions;   /* satisfying children met */
} td16o_result;

/* Depth-bounded search from (beta, omega).  Order: a node evaluates all its
 * C(n,3) position triples (i<j<k, lexicographic as N/backtracking/
 * sorting_search.cu:852-860 lays them out) before descending; a satisfying
 * child ends the search in FIRST mode (the lexicographically first one is
 * the witness); otherwise expandable children are searched depth-first in
 * the same order.  REF moves apply the reference filter
 * (sorting_search.cu:375-380, 439-441, 607-610); ALL moves take every triple
 * with a full recount.  COUNT_ALL never stops early.
 * Pinned: IDDFS distances equal BFS distances for every permutation with
 * n <= 8; ALL-move COUNT_ALL solutions equal the number of shortest sorting
 * sequences from BFS layers; expanded counts equal a naive enumeration. */
int td16o_search(const uint8_t *beta, const uint8_t *omega, int n,
                 int initialEven, float minRate, int maxDepth,
                 int fullSorting, int moveset, int mode, td16o_result *out);

/* Same search in greedy order (greedy = 1): after a node's children are
 * evaluated, the expandable ones that gain odd cycles are searched first,
 * then the others, each class in lexicographic order.  This is the kernel's
 * default FIRST order; it is this build's extension, not the reference's.
 * greedy = 0 is td16o_search.  Pinned: exhaustive searches expand the same
 * node set in either order (equal counts), found / not-found agree, and a
 * separate Python restatement gives the same witnesses on small inputs. */
int td16o_search_order(const uint8_t *beta, const uint8_t *omega, int n,
                       int initialEven, float minRate, int maxDepth,
                       int fullSorting, int moveset, int mode, int greedy,
                       td16o_result *out);

/* Same search with explicit predicate tables (the IDDFS driver uses it). */
int td16o_search_tables(const uint8_t *beta, const uint8_t *omega, int n,
                        int maxDepth, const uint32_t sat[10],
                        const uint32_t reach[10], int moveset, int mode,
                        td16o_result *out);

/* Same, giving up once more than `cap` nodes were expanded (cap 0: no
 * limit); a search that gives up reports found = -2. */
int td16o_search_capped(const uint8_t *beta, const uint8_t *omega, int n,
                        int maxDepth, const uint32_t sat[10],
                        const uint32_t reach[10], int moveset, int mode,
                        uint64_t cap, td16o_result *out);

/* Exact distance by iterative deepening: limit = max(ceil(b/3),
 * (n - c_odd)/2), +1, ... up to maxLimit.  Returns the distance or -1.
 * `out` receives the witness of the successful iteration and the sum of
 * expanded counts over all iterations. */
int td16o_iddfs(const uint8_t *beta, const uint8_t *omega, int n, int moveset,
                int maxLimit, td16o_result *out);
/* The same in either search order (greedy as in td16o_search_order). */
int td16o_iddfs_order(const uint8_t *beta, const uint8_t *omega, int n, int moveset,
                      int maxLimit, int greedy, td16o_result *out);

/* Naive enumeration: every sequence of every triple with a full recount at
 * each node, no delta and no early pruning order; counts what td16o_search
 * counts in COUNT_ALL mode.  Exponential: tiny inputs only. */
int td16o_count_naive(const uint8_t *beta, const uint8_t *omega, int n,
                      int maxDepth, const uint32_t sat[10],
                      const uint32_t reach[10], int moveset,
                      uint64_t *expanded, uint64_t *solutions);

/* Positions -> symbols by replaying block swaps on a copy of beta.
 * Follows E/base/GPUSortingSearch.java:229-239,258-264. */
int td16o_positions_to_symbols(const uint8_t *beta, int n, const int32_t *pos,
                               int len, int32_t *sym);

/* Replay of symbol triples.  Returns 0 when valid, else a negative code:
 * -1 length outside 1..maxMoves, -2 symbols not distinct / out of range,
 * -3 not an ordered triple in the current beta, -4 target not preserved,
 * -5 omega not sorted at the end, -6 malformed input.
 * Follows E/base/CertificateValidator.java:14-50 with
 * E/permutation/PermutationGroups.java:30-125 (right-to-left product) and
 * E/base/CommonOperations.java:84-108 (block swap).
 * Pinned: the reference's own fixtures CT/base/CertificateCodecTest.java:14-38,
 * ET/permutation/PermutationGroupsTest.java:9-22,
 * ET/base/CommonOperationsTest.java:21-24. */
int td16o_validate(const uint8_t *beta, const uint8_t *omega, int n,
                   const int32_t *sym, int len, int maxMoves);

/* Certificate codec (E/base/CertificateCodec.java:22-118).
 * key: omega one-line, 5 bits per image, MSB first; value: 15 bits per move.
 * Return the byte count, or -1 on invalid input. */
int td16o_codec_key(const uint8_t *omega, int n, uint8_t *out, int cap);
int td16o_codec_value(const int32_t *sym, int len, uint8_t *out, int cap);
/* Decoders return the move count / 0, or -1 when malformed (bad length or
 * nonzero padding). */
int td16o_codec_decode_value(const uint8_t *in, int nbytes, int32_t *sym, int cap);
int td16o_codec_decode_key(const uint8_t *in, int nbytes, int n, uint8_t *omega);

/* SplitMix64 (Steele, Lea, Flood 2014) and a uniform draw in [0, m) by
 * rejection of the low 2^64 mod m values. */
uint64_t td16o_splitmix64(uint64_t *state);
uint64_t td16o_uniform(uint64_t *state, uint64_t m);

/* Linear instances (nibble-packed, position p in bits 4p..4p+3; unused
 * nibbles of n < 16 hold 0xF).
 * uniform: Durstenfeld Fisher-Yates on 0..n-1 (i from n-1 down to 1,
 * j uniform in [0, i]); redrawn while the identity comes out.
 * products: n ~ U{nmin..nmax}, k ~ U{kmin..kmax}, k triples uniform over the
 * C(n+1,3) lexicographic block transpositions applied to the identity;
 * redrawn while the identity comes out.  ks receives k. */
void td16o_gen_uniform(uint64_t seed, int n, int count, uint64_t *perms);
void td16o_gen_products(uint64_t seed, int nmin, int nmax, int kmin, int kmax,
                        int count, uint8_t *ns, uint8_t *ks, uint64_t *perms);

/* Linear permutation -> the reference's (beta, omega) form with target
 * rho(x) = x+1 mod N: beta = [pi_0 .. pi_{n-1}, n], omega(y) = pred(y)+1.
 * Returns N = n+1. */
int td16o_linear_to_pair(uint64_t perm, int n, uint8_t *beta, uint8_t *omega);

/* BFS distances over all n! permutations of 0..n-1 (n <= 10) under all
 * C(n+1,3) block transpositions, indexed by lexicographic rank;
 * paths[] (optional) receives the number of shortest sorting sequences. */
int td16o_bfs(int n, uint8_t *dist, uint64_t *paths);
uint64_t td16o_perm_rank(const uint8_t *p, int n);
void td16o_perm_unrank(uint64_t r, int n, uint8_t *p);

/* Run td16o_search over a batch of pairs with `threads` POSIX threads
 * (one instance per thread at a time); used as the CPU baseline.
 * Instances are linear (ns, perms); k is the move bound per instance. */
int td16o_search_batch(int count, const uint8_t *ns, const uint64_t *perms,
                       const uint8_t *ks, int moveset, int mode, int greedy,
                       int threads, td16o_result *out);

/* The same with two more choices: iddfs = 1 computes the exact distance
 * within k by iterative deepening (found = 1, depth = distance; depth 0 for
 * a sorted input), and cap > 0 stops a search after `cap` expanded nodes
 * (found = -2: the tree is larger than the cap).  A witness that fails the
 * replay turns found into -1. */
int td16o_search_batch_ex(int count, const uint8_t *ns, const uint64_t *perms,
                          const uint8_t *ks, int moveset, int mode, int greedy, int iddfs,
                          uint64_t cap, int threads, td16o_result *out);

#ifdef __cplusplus
}
#endif
#endif
