/*
 * td16 CPU oracle -- TEST INFRASTRUCTURE ONLY (see td16_oracle.h).
 *
 * This file restates, in plain C and in the reference's own symbol form, the
 * rules of the reference's witness search so that the HIP product can be
 * checked against it.  It is deliberately written the reference's way
 * (walk-based delta, byte arrays for beta and omega, recursion) and shares no
 * code with the product.
 */
#define _POSIX_C_SOURCE 200809L
#include "td16_oracle.h"

#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Cycle algebra                                                        */
/* ------------------------------------------------------------------ */

/* N/backtracking/sorting_search.cu:211-226 */
int td16o_count_odd_cycles(const uint8_t *omega, int n) {
    uint64_t visited = 0;
    int count = 0;
    for (int s = 0; s < n; ++s) {
        if (visited & (1ULL << s)) continue;
        int size = 0, cur = s;
        do {
            visited |= 1ULL << cur;
            size++;
            cur = omega[cur];
        } while (cur != s);
        if (size & 1) count++;
    }
    return count;
}

/* Walk forward from x until one of {a, b, c} is met; returns the step count
 * and the symbol met.  N/backtracking/sorting_search.cu:246-271 */
static int walk_to_marked(const uint8_t *omega, int x, int a, int b, int c, int *hit) {
    int cur = omega[x], d = 1;
    while (cur != a && cur != b && cur != c) {
        cur = omega[cur];
        d++;
    }
    *hit = cur;
    return d;
}

/* N/backtracking/sorting_search.cu:240-332 */
int td16o_delta(const uint8_t *omega, int a, int b, int c) {
    int ta, tb, tc;
    const int da = walk_to_marked(omega, a, a, b, c, &ta);
    const int db = walk_to_marked(omega, b, a, b, c, &tb);
    const int dc = walk_to_marked(omega, c, a, b, c, &tc);

    if (ta == a && tb == b && tc == c) return TD16O_DELTA_SKIP; /* three cycles */
    if (ta == b && tb == c && tc == a) {                         /* a->b->c: split in 3 */
        const int L = da + db + dc;
        return ((da & 1) + (db & 1) + (dc & 1)) - (L & 1);
    }
    if (ta == c && tb == a && tc == b) return 0;                 /* a->c->b: intact */
    if (ta == b && tb == a) {                                    /* a,b share; c alone */
        const int L1 = da + db, L2 = dc;
        return ((da & 1) + ((db + dc) & 1)) - ((L1 & 1) + (L2 & 1));
    }
    if (tb == c && tc == b) {                                    /* b,c share; a alone */
        const int L1 = db + dc, L2 = da;
        return ((db & 1) + ((dc + da) & 1)) - ((L1 & 1) + (L2 & 1));
    }
    if (ta == c && tc == a) {                                    /* a,c share; b alone */
        const int L1 = da + dc, L2 = db;
        return ((dc & 1) + ((da + db) & 1)) - ((L1 & 1) + (L2 & 1));
    }
    return TD16O_DELTA_SKIP;
}

/* float -> int conversion with the device's semantics: truncation toward
 * zero, saturation at the int range, NaN -> 0. */
static int f2i_device(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT_MAX;
    if (x <= -2147483648.0f) return INT_MIN;
    return (int)x;
}

/* N/backtracking/sorting_search.cu:336-346 */
static int satisfies_rate(int n, int ie, float minRate, int fullSorting, int ec, int depth) {
    if (depth <= 0) return 0;
    if (fullSorting && ec != n) return 0;
    const int gained = ec - ie;
    const float achieved = (float)gained / (float)depth;
    return gained > 0 && achieved >= minRate;
}

/* N/backtracking/sorting_search.cu:348-373 */
static int can_still_reach(int n, int ie, float minRate, int maxDepth, int fullSorting,
                           int ec, int depth) {
    const int maxPossible = n - ec;
    if (fullSorting) {
        const int lb = (maxPossible + 1) / 2;
        if (depth + lb > maxDepth) return 0;
    } else {
        const int gained = ec - ie;
        const int movesLeft = maxDepth - depth;
        const int optimistic = gained + 2 * movesLeft;
        const int required = f2i_device(floorf(minRate * (float)maxDepth));
        if (optimistic < required) return 0;
        const int lb = (maxPossible + 1) / 2;
        const int best = depth + lb;
        if (best > 0) {
            const float ratio = (float)(gained + maxPossible) / (float)best;
            if (ratio < minRate) return 0;
        }
    }
    return 1;
}

void td16o_rate_tables(int n, int initialEven, float minRate, int maxDepth,
                       int fullSorting, uint32_t sat[10], uint32_t reach[10]) {
    for (int d = 0; d < 10; ++d) {
        sat[d] = 0;
        reach[d] = 0;
        for (int e = 0; e <= n && e < 32; ++e) {
            if (satisfies_rate(n, initialEven, minRate, fullSorting, e, d)) sat[d] |= 1u << e;
            if (can_still_reach(n, initialEven, minRate, maxDepth, fullSorting, e, d))
                reach[d] |= 1u << e;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Moves                                                                */
/* ------------------------------------------------------------------ */

/* Positional block swap of [i,j) and [j,k): E/base/GPUSortingSearch.java:258-264,
 * N/backtracking/transposition.cuh:5-11 */
static void swap_blocks(uint8_t *seq, int i, int j, int k) {
    uint8_t tmp[32];
    const int len = k - i;
    memcpy(tmp, seq + i, (size_t)len);
    memcpy(seq + i, tmp + (j - i), (size_t)(k - j));
    memcpy(seq + i + (k - j), tmp, (size_t)(j - i));
}

/* omega' = omega o (a b c)^-1: sorting_search.cu:443-445, 658-660 */
static void apply_symbols(uint8_t *omega, const uint8_t *parent, int a, int b, int c) {
    omega[a] = parent[c];
    omega[b] = parent[a];
    omega[c] = parent[b];
}

/* ------------------------------------------------------------------ */
/* Search                                                               */
/* ------------------------------------------------------------------ */

typedef struct {
    int n, maxDepth, moveset, mode, greedy;
    uint32_t sat[10], reach[10];
    td16o_result *out;
    int stop;
    uint64_t cap;
    int32_t path[9][3];
} search_ctx;

static void expand(search_ctx *cx, const uint8_t *beta, const uint8_t *omega, int ec, int depth) {
    const int n = cx->n;
    const int cd = depth + 1;
    uint8_t ki[680], kj[680], kk[680], kec[680];
    int nk = 0;
    uint8_t tmp[TD16O_MAX_N];

    cx->out->expanded++;
    if (cx->cap && cx->out->expanded > cx->cap) {
        cx->out->found = -2;
        cx->stop = 1;
        return;
    }
    for (int i = 0; i < n - 2; ++i) {
        for (int j = i + 1; j < n - 1; ++j) {
            for (int k = j + 1; k < n; ++k) {
                const int a = beta[i], b = beta[j], c = beta[k];
                int cec;
                if (cx->moveset == TD16O_MOVES_REF) {
                    /* sorting_search.cu:375-380, 607-610 */
                    if (omega[a] == a || omega[b] == b || omega[c] == c) continue;
                    const int d = td16o_delta(omega, a, b, c);
                    if (d != 0 && d != 2) continue;
                    cec = ec + d;
                } else {
                    memcpy(tmp, omega, (size_t)n);
                    apply_symbols(tmp, omega, a, b, c);
                    cec = td16o_count_odd_cycles(tmp, n);
                }
                if ((cx->sat[cd] >> cec) & 1u) {
                    cx->out->solutions++;
                    if (cx->mode == TD16O_FIRST) {
                        td16o_result *o = cx->out;
                        o->found = 1;
                        o->depth = cd;
                        for (int d = 0; d < depth; ++d) {
                            o->moves[d][0] = cx->path[d][0];
                            o->moves[d][1] = cx->path[d][1];
                            o->moves[d][2] = cx->path[d][2];
                        }
                        o->moves[depth][0] = i;
                        o->moves[depth][1] = j;
                        o->moves[depth][2] = k;
                        cx->stop = 1;
                        return;
                    }
                    continue;
                }
                if (cd < cx->maxDepth && ((cx->reach[cd] >> cec) & 1u)) {
                    ki[nk] = (uint8_t)i;
                    kj[nk] = (uint8_t)j;
                    kk[nk] = (uint8_t)k;
                    kec[nk] = (uint8_t)cec;
                    nk++;
                }
            }
        }
    }
    /* greedy order: children that gain odd cycles first, then the rest,
     * each class in lexicographic order (one pass in lexicographic order) */
    for (int pass = 0; pass < (cx->greedy ? 2 : 1); ++pass)
    for (int q = 0; q < nk; ++q) {
        if (cx->greedy && (pass == 0) != (kec[q] > ec)) continue;
        uint8_t cb[TD16O_MAX_N], co[TD16O_MAX_N];
        const int i = ki[q], j = kj[q], k = kk[q];
        memcpy(cb, beta, (size_t)n);
        memcpy(co, omega, (size_t)n);
        apply_symbols(co, omega, beta[i], beta[j], beta[k]);
        swap_blocks(cb, i, j, k);
        cx->path[depth][0] = i;
        cx->path[depth][1] = j;
        cx->path[depth][2] = k;
        expand(cx, cb, co, kec[q], cd);
        if (cx->stop) return;
    }
}

static int valid_pair(const uint8_t *beta, const uint8_t *omega, int n) {
    if (n < 0 || n > TD16O_MAX_N) return 0;
    uint32_t sb = 0, so = 0;
    for (int i = 0; i < n; ++i) {
        if (beta[i] >= n || omega[i] >= n) return 0;
        sb |= 1u << beta[i];
        so |= 1u << omega[i];
    }
    const uint32_t full = (n == 32) ? 0xFFFFFFFFu : ((1u << n) - 1u);
    return sb == full && so == full;
}

static int search_impl(const uint8_t *beta, const uint8_t *omega, int n, int maxDepth,
                       const uint32_t sat[10], const uint32_t reach[10], int moveset,
                       int mode, int greedy, uint64_t cap, td16o_result *out);

int td16o_search_tables(const uint8_t *beta, const uint8_t *omega, int n, int maxDepth,
                        const uint32_t sat[10], const uint32_t reach[10], int moveset,
                        int mode, td16o_result *out) {
    return search_impl(beta, omega, n, maxDepth, sat, reach, moveset, mode, 0, 0, out);
}

int td16o_search_capped(const uint8_t *beta, const uint8_t *omega, int n, int maxDepth,
                        const uint32_t sat[10], const uint32_t reach[10], int moveset,
                        int mode, uint64_t cap, td16o_result *out) {
    return search_impl(beta, omega, n, maxDepth, sat, reach, moveset, mode, 0, cap, out);
}

static int search_impl(const uint8_t *beta, const uint8_t *omega, int n, int maxDepth,
                       const uint32_t sat[10], const uint32_t reach[10], int moveset,
                       int mode, int greedy, uint64_t cap, td16o_result *out) {
    memset(out, 0, sizeof(*out));
    if (!valid_pair(beta, omega, n)) return -1;
    if (maxDepth < 1 || maxDepth > TD16O_MAX_DEPTH) return -1;
    if (n < 3) return 0; /* C(n,3) == 0: sorting_search.cu:826-827 */
    search_ctx cx;
    memset(&cx, 0, sizeof(cx));
    cx.n = n;
    cx.maxDepth = maxDepth;
    cx.moveset = moveset;
    cx.mode = mode;
    cx.greedy = greedy;
    memcpy(cx.sat, sat, sizeof(cx.sat));
    memcpy(cx.reach, reach, sizeof(cx.reach));
    cx.out = out;
    cx.cap = cap;
    expand(&cx, beta, omega, td16o_count_odd_cycles(omega, n), 0);
    return 0;
}

int td16o_search(const uint8_t *beta, const uint8_t *omega, int n, int initialEven,
                 float minRate, int maxDepth, int fullSorting, int moveset, int mode,
                 td16o_result *out) {
    uint32_t sat[10], reach[10];
    td16o_rate_tables(n, initialEven, minRate, maxDepth, fullSorting, sat, reach);
    return td16o_search_tables(beta, omega, n, maxDepth, sat, reach, moveset, mode, out);
}

int td16o_search_order(const uint8_t *beta, const uint8_t *omega, int n, int initialEven,
                       float minRate, int maxDepth, int fullSorting, int moveset, int mode,
                       int greedy, td16o_result *out) {
    uint32_t sat[10], reach[10];
    td16o_rate_tables(n, initialEven, minRate, maxDepth, fullSorting, sat, reach);
    return search_impl(beta, omega, n, maxDepth, sat, reach, moveset, mode, greedy, 0, out);
}

int td16o_iddfs(const uint8_t *beta, const uint8_t *omega, int n, int moveset, int maxLimit,
                td16o_result *out) {
    return td16o_iddfs_order(beta, omega, n, moveset, maxLimit, 0, out);
}

int td16o_iddfs_order(const uint8_t *beta, const uint8_t *omega, int n, int moveset,
                      int maxLimit, int greedy, td16o_result *out) {
    memset(out, 0, sizeof(*out));
    if (!valid_pair(beta, omega, n)) return -1;
    const int odd = td16o_count_odd_cycles(omega, n);
    if (odd == n) return 0;
    int b = 0;
    for (int s = 0; s < n; ++s) b += omega[s] != s;
    int limit = (n - odd) / 2;
    if ((b + 2) / 3 > limit) limit = (b + 2) / 3;
    if (limit < 1) limit = 1;
    uint64_t total = 0;
    for (; limit <= maxLimit; ++limit) {
        uint32_t sat[10], reach[10];
        td16o_rate_tables(n, odd, 0.0f, limit, 1, sat, reach);
        td16o_result r;
        search_impl(beta, omega, n, limit, sat, reach, moveset, TD16O_FIRST, greedy, 0, &r);
        total += r.expanded;
        if (r.found) {
            *out = r;
            out->expanded = total;
            return limit;
        }
    }
    out->expanded = total;
    return -1;
}

/* Cycle labels of omega, by walking: lab[s] = smallest symbol of s's cycle. */
static void cycle_labels(const uint8_t *omega, int n, uint8_t *lab) {
    for (int s = 0; s < n; ++s) lab[s] = 0xFF;
    for (int s = 0; s < n; ++s) {
        if (lab[s] != 0xFF) continue;
        int cur = s;
        do {
            lab[cur] = (uint8_t)s;
            cur = omega[cur];
        } while (cur != s);
    }
}

typedef struct {
    int n, maxDepth, moveset;
    const uint32_t *sat, *reach;
    uint64_t expanded, solutions;
} naive_ctx;

static void naive(naive_ctx *cx, const uint8_t *beta, const uint8_t *omega, int depth) {
    const int n = cx->n;
    uint8_t lab[TD16O_MAX_N];
    cx->expanded++;
    cycle_labels(omega, n, lab);
    const int ec = td16o_count_odd_cycles(omega, n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                if (!(i < j && j < k)) continue;
                uint8_t cb[TD16O_MAX_N], co[TD16O_MAX_N];
                const int a = beta[i], b = beta[j], c = beta[k];
                memcpy(cb, beta, (size_t)n);
                memcpy(co, omega, (size_t)n);
                co[a] = omega[c];
                co[b] = omega[a];
                co[c] = omega[b];
                swap_blocks(cb, i, j, k);
                const int cec = td16o_count_odd_cycles(co, n);
                if (cx->moveset == TD16O_MOVES_REF) {
                    if (omega[a] == a || omega[b] == b || omega[c] == c) continue;
                    if (lab[a] != lab[b] && lab[b] != lab[c] && lab[a] != lab[c]) continue;
                    if (cec - ec != 0 && cec - ec != 2) continue;
                }
                if ((cx->sat[depth + 1] >> cec) & 1u) {
                    cx->solutions++;
                    continue;
                }
                if (depth + 1 < cx->maxDepth && ((cx->reach[depth + 1] >> cec) & 1u))
                    naive(cx, cb, co, depth + 1);
            }
}

int td16o_count_naive(const uint8_t *beta, const uint8_t *omega, int n, int maxDepth,
                      const uint32_t sat[10], const uint32_t reach[10], int moveset,
                      uint64_t *expanded, uint64_t *solutions) {
    *expanded = 0;
    *solutions = 0;
    if (!valid_pair(beta, omega, n) || maxDepth < 1 || maxDepth > TD16O_MAX_DEPTH) return -1;
    if (n < 3) return 0;
    naive_ctx cx = {n, maxDepth, moveset, sat, reach, 0, 0};
    naive(&cx, beta, omega, 0);
    *expanded = cx.expanded;
    *solutions = cx.solutions;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Caller-side replay                                                   */
/* ------------------------------------------------------------------ */

int td16o_positions_to_symbols(const uint8_t *beta, int n, const int32_t *pos, int len,
                               int32_t *sym) {
    uint8_t cur[TD16O_MAX_N];
    if (n < 0 || n > TD16O_MAX_N) return -1;
    memcpy(cur, beta, (size_t)n);
    for (int m = 0; m < len; ++m) {
        const int i = pos[3 * m], j = pos[3 * m + 1], k = pos[3 * m + 2];
        if (!(0 <= i && i < j && j < k && k < n)) return -1;
        sym[3 * m] = cur[i];
        sym[3 * m + 1] = cur[j];
        sym[3 * m + 2] = cur[k];
        swap_blocks(cur, i, j, k);
    }
    return 0;
}

/* compose(f, g)[x] = f[g[x]]: the right-to-left product of
 * E/permutation/PermutationGroups.java:30-125 (last factor applied first). */
static void compose(const uint8_t *f, const uint8_t *g, int n, uint8_t *out) {
    for (int x = 0; x < n; ++x) out[x] = f[g[x]];
}

/* Function form of a cyclic sequence: next[seq[p]] = seq[p+1 mod n]. */
static void cycle_function(const uint8_t *seq, int n, uint8_t *fn) {
    for (int p = 0; p < n; ++p) fn[seq[p]] = seq[(p + 1) % n];
}

int td16o_validate(const uint8_t *beta0, const uint8_t *omega0, int n, const int32_t *sym,
                   int len, int maxMoves) {
    /* CertificateValidator.java:16-19 */
    if (len < 1 || len > maxMoves) return -1;
    if (!valid_pair(beta0, omega0, n)) return -6;
    uint8_t beta[TD16O_MAX_N], omega[TD16O_MAX_N], bfun[TD16O_MAX_N], rho[TD16O_MAX_N];
    memcpy(beta, beta0, (size_t)n);
    memcpy(omega, omega0, (size_t)n);
    cycle_function(beta, n, bfun);
    compose(omega, bfun, n, rho); /* rho = omega . beta */
    for (int m = 0; m < len; ++m) {
        const int a = sym[3 * m], b = sym[3 * m + 1], c = sym[3 * m + 2];
        /* CertificateValidator.java:39-50 */
        if (a == b || b == c || a == c) return -2;
        if (a < 0 || b < 0 || c < 0 || a >= n || b >= n || c >= n) return -2;
        int pa = -1, pb = -1, pc = -1;
        for (int p = 0; p < n; ++p) {
            if (beta[p] == a) pa = p;
            if (beta[p] == b) pb = p;
            if (beta[p] == c) pc = p;
        }
        if (pa < 0 || pb < 0 || pc < 0 || !(pa < pb && pb < pc)) return -3;
        /* omega <- omega . (a b c)^-1 ; the inverse cycle maps a->c, b->a, c->b */
        uint8_t inv[TD16O_MAX_N], next[TD16O_MAX_N];
        for (int x = 0; x < n; ++x) inv[x] = (uint8_t)x;
        inv[a] = (uint8_t)c;
        inv[b] = (uint8_t)a;
        inv[c] = (uint8_t)b;
        compose(omega, inv, n, next);
        memcpy(omega, next, (size_t)n);
        /* CommonOperations.applyTranspositionOptimized: sorted indexes, block swap */
        swap_blocks(beta, pa, pb, pc);
        cycle_function(beta, n, bfun);
        uint8_t tgt[TD16O_MAX_N];
        compose(omega, bfun, n, tgt);
        if (memcmp(tgt, rho, (size_t)n) != 0) return -4;
    }
    for (int x = 0; x < n; ++x)
        if (omega[x] != x) return -5;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Codec: E/base/CertificateCodec.java:22-118                            */
/* ------------------------------------------------------------------ */

static int write5(uint8_t *buf, int bit, int v) {
    for (int q = 0; q < 5; ++q) {
        const int vb = (v >> (4 - q)) & 1;
        const int ob = bit + q;
        buf[ob / 8] |= (uint8_t)(vb << (7 - ob % 8));
    }
    return bit + 5;
}

static int read5(const uint8_t *buf, int bit) {
    int v = 0;
    for (int q = 0; q < 5; ++q) {
        const int ib = bit + q;
        v = (v << 1) | ((buf[ib / 8] >> (7 - ib % 8)) & 1);
    }
    return v;
}

static int payload_ok(const uint8_t *buf, int nbytes, long bits) {
    if (nbytes != (int)((bits + 7) / 8)) return 0;
    const int pad = (int)(8L * nbytes - bits);
    if (pad != 0 && (buf[nbytes - 1] & ((1 << pad) - 1)) != 0) return 0;
    return 1;
}

int td16o_codec_key(const uint8_t *omega, int n, uint8_t *out, int cap) {
    if (n <= 0 || n > 32) return -1;
    const int nb = (5 * n + 7) / 8;
    if (nb > cap) return -1;
    memset(out, 0, (size_t)nb);
    int bit = 0;
    for (int s = 0; s < n; ++s) {
        if (omega[s] >= 32) return -1;
        bit = write5(out, bit, omega[s]);
    }
    return nb;
}

int td16o_codec_value(const int32_t *sym, int len, uint8_t *out, int cap) {
    if (len <= 0) return -1;
    const int nb = (15 * len + 7) / 8;
    if (nb > cap) return -1;
    memset(out, 0, (size_t)nb);
    int bit = 0;
    for (int m = 0; m < 3 * len; ++m) {
        if (sym[m] < 0 || sym[m] >= 32) return -1;
        bit = write5(out, bit, sym[m]);
    }
    return nb;
}

int td16o_codec_decode_value(const uint8_t *in, int nbytes, int32_t *sym, int cap) {
    if (in == NULL || nbytes <= 0) return -1;
    const int moves = (int)((8L * nbytes) / 15);
    if (moves == 0) return -1;
    if (!payload_ok(in, nbytes, 15L * moves)) return -1;
    if (moves > cap) return -1;
    for (int m = 0; m < 3 * moves; ++m) sym[m] = read5(in, 5 * m);
    return moves;
}

int td16o_codec_decode_key(const uint8_t *in, int nbytes, int n, uint8_t *omega) {
    if (n <= 0 || n > 32) return -1;
    if (!payload_ok(in, nbytes, 5L * n)) return -1;
    for (int s = 0; s < n; ++s) omega[s] = (uint8_t)read5(in, 5 * s);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Instances                                                            */
/* ------------------------------------------------------------------ */

uint64_t td16o_splitmix64(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t td16o_uniform(uint64_t *state, uint64_t m) {
    const uint64_t threshold = (0 - m) % m; /* 2^64 mod m */
    for (;;) {
        const uint64_t r = td16o_splitmix64(state);
        if (r >= threshold) return r % m;
    }
}

static uint64_t pack_nibbles(const uint8_t *p, int n) {
    uint64_t v = ~0ULL;
    for (int q = 0; q < n; ++q) {
        v &= ~(0xFULL << (4 * q));
        v |= (uint64_t)p[q] << (4 * q);
    }
    return v;
}

static int is_identity(const uint8_t *p, int n) {
    for (int q = 0; q < n; ++q)
        if (p[q] != q) return 0;
    return 1;
}

void td16o_gen_uniform(uint64_t seed, int n, int count, uint64_t *perms) {
    uint64_t st = seed;
    for (int c = 0; c < count; ++c) {
        uint8_t p[16];
        do {
            for (int q = 0; q < n; ++q) p[q] = (uint8_t)q;
            for (int i = n - 1; i >= 1; --i) {
                const int j = (int)td16o_uniform(&st, (uint64_t)i + 1);
                const uint8_t t = p[i];
                p[i] = p[j];
                p[j] = t;
            }
        } while (is_identity(p, n));
        perms[c] = pack_nibbles(p, n);
    }
}

/* t-th triple (i<j<k) of 0..N-1 in lexicographic order */
static void lex_unrank3(uint64_t t, int N, int *i, int *j, int *k) {
    for (int a = 0; a < N - 2; ++a)
        for (int b = a + 1; b < N - 1; ++b) {
            const uint64_t cnt = (uint64_t)(N - 1 - b);
            if (t < cnt) {
                *i = a;
                *j = b;
                *k = b + 1 + (int)t;
                return;
            }
            t -= cnt;
        }
    *i = *j = *k = -1;
}

void td16o_gen_products(uint64_t seed, int nmin, int nmax, int kmin, int kmax, int count,
                        uint8_t *ns, uint8_t *ks, uint64_t *perms) {
    uint64_t st = seed;
    for (int c = 0; c < count; ++c) {
        uint8_t p[17];
        int n, k;
        do {
            n = nmin + (int)td16o_uniform(&st, (uint64_t)(nmax - nmin + 1));
            k = kmin + (int)td16o_uniform(&st, (uint64_t)(kmax - kmin + 1));
            for (int q = 0; q < n; ++q) p[q] = (uint8_t)q;
            const int N = n + 1;
            const uint64_t T = (uint64_t)N * (N - 1) * (N - 2) / 6;
            for (int m = 0; m < k; ++m) {
                int i, j, kk;
                lex_unrank3(td16o_uniform(&st, T), N, &i, &j, &kk);
                swap_blocks(p, i, j, kk);
            }
        } while (is_identity(p, n));
        ns[c] = (uint8_t)n;
        ks[c] = (uint8_t)k;
        perms[c] = pack_nibbles(p, n);
    }
}

int td16o_linear_to_pair(uint64_t perm, int n, uint8_t *beta, uint8_t *omega) {
    const int N = n + 1;
    for (int q = 0; q < n; ++q) beta[q] = (uint8_t)((perm >> (4 * q)) & 0xF);
    beta[n] = (uint8_t)n;
    for (int p = 0; p < N; ++p) {
        const int prev = beta[(p + N - 1) % N];
        omega[beta[p]] = (uint8_t)((prev + 1) % N);
    }
    return N;
}

/* ------------------------------------------------------------------ */
/* BFS over S_n                                                         */
/* ------------------------------------------------------------------ */

uint64_t td16o_perm_rank(const uint8_t *p, int n) {
    uint64_t r = 0;
    for (int i = 0; i < n; ++i) {
        int smaller = 0;
        for (int j = i + 1; j < n; ++j) smaller += p[j] < p[i];
        r = r * (uint64_t)(n - i) + (uint64_t)smaller;
    }
    return r;
}

void td16o_perm_unrank(uint64_t r, int n, uint8_t *p) {
    int digits[16];
    for (int i = n - 1; i >= 0; --i) {
        const uint64_t base = (uint64_t)(n - i);
        digits[i] = (int)(r % base);
        r /= base;
    }
    int used[16] = {0};
    for (int i = 0; i < n; ++i) {
        int c = digits[i];
        for (int v = 0; v < n; ++v) {
            if (used[v]) continue;
            if (c == 0) {
                p[i] = (uint8_t)v;
                used[v] = 1;
                break;
            }
            c--;
        }
    }
}

int td16o_bfs(int n, uint8_t *dist, uint64_t *paths) {
    if (n < 1 || n > 10) return -1;
    uint64_t total = 1;
    for (int q = 2; q <= n; ++q) total *= (uint64_t)q;
    memset(dist, 0xFF, (size_t)total);
    if (paths) memset(paths, 0, sizeof(uint64_t) * total);
    uint64_t *frontier = (uint64_t *)malloc(sizeof(uint64_t) * total);
    uint64_t *nextf = (uint64_t *)malloc(sizeof(uint64_t) * total);
    if (!frontier || !nextf) {
        free(frontier);
        free(nextf);
        return -1;
    }
    uint8_t id[16];
    for (int q = 0; q < n; ++q) id[q] = (uint8_t)q;
    const uint64_t r0 = td16o_perm_rank(id, n);
    dist[r0] = 0;
    if (paths) paths[r0] = 1;
    uint64_t fsz = 1;
    frontier[0] = r0;
    int level = 0;
    while (fsz) {
        uint64_t nsz = 0;
        for (uint64_t f = 0; f < fsz; ++f) {
            uint8_t p[16], q[16];
            td16o_perm_unrank(frontier[f], n, p);
            for (int i = 0; i < n; ++i)
                for (int j = i + 1; j < n; ++j)
                    for (int k = j + 1; k <= n; ++k) {
                        memcpy(q, p, (size_t)n);
                        swap_blocks(q, i, j, k);
                        const uint64_t r = td16o_perm_rank(q, n);
                        if (dist[r] == 0xFF) {
                            dist[r] = (uint8_t)(level + 1);
                            nextf[nsz++] = r;
                        }
                        if (paths && dist[r] == level + 1) paths[r] += paths[frontier[f]];
                    }
        }
        uint64_t *t = frontier;
        frontier = nextf;
        nextf = t;
        fsz = nsz;
        level++;
    }
    free(frontier);
    free(nextf);
    return level - 1;
}

/* ------------------------------------------------------------------ */
/* Threaded batch (CPU baseline)                                        */
/* ------------------------------------------------------------------ */

typedef struct {
    int count;
    const uint8_t *ns;
    const uint64_t *perms;
    const uint8_t *ks;
    int moveset, mode, greedy, iddfs;
    uint64_t cap;
    td16o_result *out;
    int next;
    pthread_mutex_t mu;
} batch_ctx;

static void *batch_worker(void *arg) {
    batch_ctx *b = (batch_ctx *)arg;
    for (;;) {
        pthread_mutex_lock(&b->mu);
        const int idx = b->next++;
        pthread_mutex_unlock(&b->mu);
        if (idx >= b->count) break;
        uint8_t beta[TD16O_MAX_N], omega[TD16O_MAX_N];
        const int n = b->ns[idx];
        const int N = td16o_linear_to_pair(b->perms[idx], n, beta, omega);
        const int odd = td16o_count_odd_cycles(omega, N);
        const int target = (N - odd) / 2;
        const int k = b->ks[idx];
        /* caller convention E/UnorientedCycles.java:34 */
        const float minRate = 2.0f / ((float)k / (float)target);
        td16o_result *r = &b->out[idx];
        if (b->iddfs) {
            /* exact distance within k (depth 0 for a sorted input) */
            td16o_iddfs_order(beta, omega, N, b->moveset, k, b->greedy, r);
            if (odd == N) r->found = 1;
        } else {
            uint32_t sat[10], reach[10];
            td16o_rate_tables(N, odd, minRate, k, 1, sat, reach);
            search_impl(beta, omega, N, k, sat, reach, b->moveset, b->mode, b->greedy, b->cap, r);
        }
        if (r->found == 1 && r->depth > 0) {
            int32_t sym[27];
            td16o_positions_to_symbols(beta, N, &r->moves[0][0], r->depth, sym);
            if (td16o_validate(beta, omega, N, sym, r->depth, k) != 0) r->found = -1;
        }
    }
    return NULL;
}

int td16o_search_batch(int count, const uint8_t *ns, const uint64_t *perms, const uint8_t *ks,
                       int moveset, int mode, int greedy, int threads, td16o_result *out) {
    return td16o_search_batch_ex(count, ns, perms, ks, moveset, mode, greedy, 0, 0, threads,
                                 out);
}

int td16o_search_batch_ex(int count, const uint8_t *ns, const uint64_t *perms,
                          const uint8_t *ks, int moveset, int mode, int greedy, int iddfs,
                          uint64_t cap, int threads, td16o_result *out) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    batch_ctx b;
    b.count = count;
    b.ns = ns;
    b.perms = perms;
    b.ks = ks;
    b.moveset = moveset;
    b.mode = mode;
    b.greedy = greedy;
    b.iddfs = iddfs;
    b.cap = cap;
    b.out = out;
    b.next = 0;
    pthread_mutex_init(&b.mu, NULL);
    pthread_t tid[256];
    for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, batch_worker, &b);
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    pthread_mutex_destroy(&b.mu);
    return 0;
}
