/*
 * egtb_oracle.c -- CPU restatement of the reference's classic tablebase
 * builder (forward layer enumeration, dedup, backward success-rate solve).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product path never calls it.
 *
 * Every function cites the reference behaviour it restates.  Pinning (how a
 * plausible mistake in each function is caught) is listed in tests/:
 *   - row move table: exhaustive 65536-row check against an independent
 *     pure-Python merge rule (tests/test_oracle_primitives.py);
 *   - transpose / symmetries / canonical forms: group laws and numpy
 *     rot90/flip restatement on random boards;
 *   - predicates: truth tables on the catalogue masks;
 *   - forward sets and backward rates: per-layer SHA-256 digests and sampled
 *     records produced by the reference's own C++ compiled in this container
 *     (oracle/ref), plus exact-rational brute force on tiny patterns.
 *
 * Board encoding (native_core/include/BoardCodec.h:20-29): nibble k holds
 * the cell at row 3-k/4, column 3-k%4; value v is tile 2^v, 0 is empty.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <omp.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* Row move table                                                      */
/* ------------------------------------------------------------------ */

/* Result row after sliding a 4-cell row toward its high nibble ("left").
 * Restates StandardMergePolicy::_internal_merge (BoardMover.h:7-55): slide
 * non-empty tiles, merge equal neighbours once from the move side, a 32768
 * tile (nibble 0xF) never merges. */
static uint16_t slide_row_high(uint16_t row) {
    unsigned tiles[4], n = 0, out[4] = {0, 0, 0, 0}, m = 0;
    for (int c = 0; c < 4; ++c) {
        unsigned v = (row >> (12 - 4 * c)) & 0xFu;
        if (v) tiles[n++] = v;
    }
    for (unsigned i = 0; i < n; ++i) {
        if (i + 1 < n && tiles[i] == tiles[i + 1] && tiles[i] != 0xFu) {
            out[m++] = tiles[i] + 1;
            ++i;
        } else {
            out[m++] = tiles[i];
        }
    }
    return (uint16_t)((out[0] << 12) | (out[1] << 8) | (out[2] << 4) | out[3]);
}

static uint16_t rev_row(uint16_t r) {
    return (uint16_t)(((r & 0xFu) << 12) | ((r & 0xF0u) << 4) | ((r >> 4) & 0xF0u) | (r >> 12));
}

static uint16_t LEFT_ROW[65536];
static uint16_t RIGHT_ROW[65536];
static int g_ready = 0;

ORC_API void orc_init(void) {
    if (g_ready) return;
    for (uint32_t r = 0; r < 65536; ++r) {
        LEFT_ROW[r] = slide_row_high((uint16_t)r);
        /* right = reverse, slide, reverse (BoardMover.h:16,42) */
        RIGHT_ROW[r] = rev_row(slide_row_high(rev_row((uint16_t)r)));
    }
    g_ready = 1;
}

ORC_API uint16_t orc_row_left(uint16_t r) { orc_init(); return LEFT_ROW[r]; }
ORC_API uint16_t orc_row_right(uint16_t r) { orc_init(); return RIGHT_ROW[r]; }

/* ------------------------------------------------------------------ */
/* Board transforms                                                    */
/* ------------------------------------------------------------------ */

/* Main-diagonal transpose of the 4x4 nibble matrix, as two delta swaps
 * (2x2 blocks, then cells inside blocks).  Same map as reverse_board /
 * ReverseUL (CommonMover.h:33-37, Calculator.h:40-44). */
static inline uint64_t transpose(uint64_t x) {
    uint64_t t = ((x >> 24) ^ x) & 0x00000000FF00FF00ull;
    x ^= t ^ (t << 24);
    t = ((x >> 12) ^ x) & 0x0000F0F00000F0F0ull;
    x ^= t ^ (t << 12);
    return x;
}

/* Mirror columns (Calculator::ReverseLR, Calculator.h:28-32). */
static inline uint64_t mirror_lr(uint64_t x) {
    x = ((x >> 8) & 0x00FF00FF00FF00FFull) | ((x & 0x00FF00FF00FF00FFull) << 8);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
    return x;
}

/* Mirror rows (Calculator::ReverseUD, Calculator.h:34-38). */
static inline uint64_t mirror_ud(uint64_t x) {
    x = (x >> 32) | (x << 32);
    x = ((x >> 16) & 0x0000FFFF0000FFFFull) | ((x & 0x0000FFFF0000FFFFull) << 16);
    return x;
}

static inline uint64_t rows_apply(uint64_t b, const uint16_t *tab) {
    return (uint64_t)tab[b & 0xFFFF] | ((uint64_t)tab[(b >> 16) & 0xFFFF] << 16) |
           ((uint64_t)tab[(b >> 32) & 0xFFFF] << 32) | ((uint64_t)tab[b >> 48] << 48);
}

/* The four moves in the reference's move_all_dir order: left, right, up,
 * down (CommonMover.h:136-139); up/down act on the transpose. */
static inline void move4(uint64_t b, uint64_t out[4]) {
    uint64_t t = transpose(b);
    out[0] = rows_apply(b, LEFT_ROW);
    out[1] = rows_apply(b, RIGHT_ROW);
    out[2] = transpose(rows_apply(t, LEFT_ROW));
    out[3] = transpose(rows_apply(t, RIGHT_ROW));
}

ORC_API uint64_t orc_move(uint64_t b, int dir) {
    uint64_t o[4];
    orc_init();
    move4(b, o);
    return o[dir & 3];
}

ORC_API uint64_t orc_transpose(uint64_t b) { return transpose(b); }
ORC_API uint64_t orc_mirror_lr(uint64_t b) { return mirror_lr(b); }
ORC_API uint64_t orc_mirror_ud(uint64_t b) { return mirror_ud(b); }

/* Canonical representative for the pattern's symmetry mode
 * (apply_canonical, BookGenerator.cpp:231-249; Calculator.h:80-98):
 * 0 identity, 1 full (min over the 8 dihedral images), 2 diagonal
 * (min with the transpose), 3 horizontal (min with the column mirror). */
static inline uint64_t umin(uint64_t a, uint64_t b) { return a < b ? a : b; }

static inline uint64_t canonical(uint64_t b, int mode) {
    switch (mode) {
    case 1: {
        uint64_t t = transpose(b);
        uint64_t m = umin(b, t);
        m = umin(m, mirror_lr(b));
        m = umin(m, mirror_ud(b));
        m = umin(m, mirror_lr(mirror_ud(b)));
        m = umin(m, mirror_lr(t));
        m = umin(m, mirror_ud(t));
        m = umin(m, mirror_lr(mirror_ud(t)));
        return m;
    }
    case 2: return umin(b, transpose(b));
    case 3: return umin(b, mirror_lr(b));
    default: return b;
    }
}

ORC_API uint64_t orc_canonical(uint64_t b, int mode) { return canonical(b, mode); }

/* ------------------------------------------------------------------ */
/* Pattern spec and predicates                                         */
/* ------------------------------------------------------------------ */

typedef struct {
    uint64_t masks[16];
    int n_masks;
    uint8_t shifts[16];
    int n_shifts;
    int symm;
    int target; /* log2 of the target tile */
} orc_spec;

/* is_pattern (Formation.h:11-17): no masks means every board matches. */
static inline int is_pattern(uint64_t b, const orc_spec *s) {
    if (s->n_masks == 0) return 1;
    for (int i = 0; i < s->n_masks; ++i)
        if ((b & s->masks[i]) == s->masks[i]) return 1;
    return 0;
}

/* is_success_by_shifts (Formation.h:32-43): no shifts means success. */
static inline int is_success(uint64_t b, const orc_spec *s) {
    if (s->n_shifts == 0) return 1;
    for (int i = 0; i < s->n_shifts; ++i)
        if (((b >> s->shifts[i]) & 0xFu) == (uint64_t)s->target) return 1;
    return 0;
}

static void make_spec(orc_spec *s, const uint64_t *masks, int n_masks, const uint8_t *shifts,
                      int n_shifts, int symm, int target) {
    memset(s, 0, sizeof(*s));
    s->n_masks = n_masks > 16 ? 16 : n_masks;
    for (int i = 0; i < s->n_masks; ++i) s->masks[i] = masks[i];
    s->n_shifts = n_shifts > 16 ? 16 : n_shifts;
    for (int i = 0; i < s->n_shifts; ++i) s->shifts[i] = shifts[i];
    s->symm = symm;
    s->target = target;
}

ORC_API int orc_is_pattern(uint64_t b, const uint64_t *masks, int n_masks) {
    orc_spec s;
    make_spec(&s, masks, n_masks, NULL, 0, 0, 0);
    return is_pattern(b, &s);
}

ORC_API int orc_is_success(uint64_t b, int target, const uint8_t *shifts, int n_shifts) {
    orc_spec s;
    make_spec(&s, NULL, 0, shifts, n_shifts, 0, target);
    return is_success(b, &s);
}

/* ------------------------------------------------------------------ */
/* Growable u64 vector                                                  */
/* ------------------------------------------------------------------ */

typedef struct {
    uint64_t *p;
    size_t n, cap;
} vec64;

static void v_push(vec64 *v, uint64_t x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 1024;
        v->p = (uint64_t *)realloc(v->p, v->cap * sizeof(uint64_t));
        if (!v->p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    }
    v->p[v->n++] = x;
}

/* ------------------------------------------------------------------ */
/* Parallel LSD radix sort + unique (stands in for sort_array and      */
/* parallel_unique, BookGeneratorUtils.cpp:138-206)                     */
/* ------------------------------------------------------------------ */

static void radix_sort(uint64_t *a, size_t n, int nthreads) {
    if (n < 2) return;
    uint64_t *tmp = (uint64_t *)malloc(n * sizeof(uint64_t));
    if (!tmp) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    uint64_t diff = 0, first = a[0];
    for (size_t i = 1; i < n; ++i) diff |= a[i] ^ first;
    size_t *hist = (size_t *)calloc((size_t)nthreads * 256, sizeof(size_t));
    uint64_t *src = a, *dst = tmp;
    for (int shift = 0; shift < 64; shift += 8) {
        if (((diff >> shift) & 0xFF) == 0) continue;
        memset(hist, 0, (size_t)nthreads * 256 * sizeof(size_t));
#pragma omp parallel num_threads(nthreads)
        {
            int t = omp_get_thread_num();
            size_t lo = n * (size_t)t / (size_t)nthreads, hi = n * (size_t)(t + 1) / (size_t)nthreads;
            size_t *h = hist + (size_t)t * 256;
            for (size_t i = lo; i < hi; ++i) h[(src[i] >> shift) & 0xFF]++;
#pragma omp barrier
#pragma omp single
            {
                size_t sum = 0;
                for (int d = 0; d < 256; ++d)
                    for (int tt = 0; tt < nthreads; ++tt) {
                        size_t c = hist[(size_t)tt * 256 + d];
                        hist[(size_t)tt * 256 + d] = sum;
                        sum += c;
                    }
            }
            for (size_t i = lo; i < hi; ++i) dst[h[(src[i] >> shift) & 0xFF]++] = src[i];
        }
        uint64_t *sw = src; src = dst; dst = sw;
    }
    if (src != a) memcpy(a, src, n * sizeof(uint64_t));
    free(hist);
    free(tmp);
}

static size_t unique_sorted(uint64_t *a, size_t n) {
    if (n < 2) return n;
    size_t w = 1;
    for (size_t i = 1; i < n; ++i)
        if (a[i] != a[w - 1]) a[w++] = a[i];
    return w;
}

ORC_API size_t orc_sort_unique(uint64_t *a, size_t n, int nthreads) {
    radix_sort(a, n, nthreads > 0 ? nthreads : 1);
    return unique_sorted(a, n);
}

/* ------------------------------------------------------------------ */
/* Forward expansion (gen_boards_simple / _naive, BookGenerator.cpp:   */
/* 473-589, 718-760)                                                    */
/* ------------------------------------------------------------------ */

/* For every parent not skipped by the success check, each empty nibble
 * k = 0..15, spawn 2 (nibble 1) and spawn 4 (nibble 2), the four moves;
 * keep moved != spawned && is_pattern(moved), store canonical(moved).
 * Spawn-2 children go to out2, spawn-4 children to out4 (unordered). */
static void expand(const uint64_t *par, size_t n, const orc_spec *s, int do_check, vec64 *out2,
                   vec64 *out4, int nthreads) {
    vec64 *loc2 = (vec64 *)calloc((size_t)nthreads, sizeof(vec64));
    vec64 *loc4 = (vec64 *)calloc((size_t)nthreads, sizeof(vec64));
#pragma omp parallel num_threads(nthreads)
    {
        int t = omp_get_thread_num();
#pragma omp for schedule(dynamic, 1024)
        for (size_t i = 0; i < n; ++i) {
            uint64_t b = par[i];
            if (do_check && is_success(b, s)) continue;
            for (int k = 0; k < 16; ++k) {
                if ((b >> (4 * k)) & 0xFu) continue;
                for (int sp = 1; sp <= 2; ++sp) {
                    uint64_t spawned = b | ((uint64_t)sp << (4 * k));
                    uint64_t mv[4];
                    move4(spawned, mv);
                    for (int d = 0; d < 4; ++d) {
                        if (mv[d] != spawned && is_pattern(mv[d], s))
                            v_push(sp == 1 ? &loc2[t] : &loc4[t], canonical(mv[d], s->symm));
                    }
                }
            }
        }
    }
    for (int t = 0; t < nthreads; ++t) {
        for (size_t i = 0; i < loc2[t].n; ++i) v_push(out2, loc2[t].p[i]);
        for (size_t i = 0; i < loc4[t].n; ++i) v_push(out4, loc4[t].p[i]);
        free(loc2[t].p);
        free(loc4[t].p);
    }
    free(loc2);
    free(loc4);
}

/* ------------------------------------------------------------------ */
/* Backward solve (recalculate_layer, BookSolver.cpp:413-470)          */
/* ------------------------------------------------------------------ */

typedef struct {
    uint64_t *keys; /* ascending except layer 0 (seed order) */
    float *rates;
    size_t n;
} orc_book;

/* Exact-match lookup in a sorted layer; absent -> 0 (search_arr,
 * BookSolver.cpp:317-327 with HybridSearch::exact_search). */
static inline float lookup(const orc_book *L, uint64_t key) {
    size_t lo = 0, hi = L->n;
    while (lo < hi) {
        size_t mid = lo + ((hi - lo) >> 1);
        if (L->keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    return (lo < L->n && L->keys[lo] == key) ? L->rates[lo] : 0.0f;
}

/* One state's rate.  FP order (BookSolver.cpp:436-468): sp in double, per
 * empty nibble ascending: sp += (double)best2 * (1 - p4), then
 * sp += (double)best4 * p4; rate = (float)(sp / E).  best_x is the float
 * max over valid moves of the looked-up rate (0 when no valid move).
 * fma_mode 1 evaluates each update as one fused multiply-add, which is what
 * the reference's -Ofast -march=x86-64-v3 build does (CMakeLists.txt:82-83);
 * fma_mode 0 rounds the product first. */
static inline float solve_state(uint64_t b, const orc_book *L1, const orc_book *L2,
                                const orc_spec *s, int do_check, double p4, int fma_mode) {
    if (do_check && is_success(b, s)) return 1.0f;
    const double w2 = 1.0 - p4;
    double sp = 0.0;
    int empty = 0;
    for (int k = 0; k < 16; ++k) {
        if ((b >> (4 * k)) & 0xFu) continue;
        ++empty;
        float best[2] = {0.0f, 0.0f};
        for (int si = 0; si < 2; ++si) {
            uint64_t spawned = b | ((uint64_t)(si + 1) << (4 * k));
            uint64_t mv[4];
            move4(spawned, mv);
            for (int d = 0; d < 4; ++d) {
                if (mv[d] != spawned && is_pattern(mv[d], s)) {
                    float r = lookup(si == 0 ? L1 : L2, canonical(mv[d], s->symm));
                    best[si] = best[si] < r ? r : best[si];
                }
            }
        }
        if (fma_mode) {
            sp = fma((double)best[0], w2, sp);
            sp = fma((double)best[1], p4, sp);
        } else {
            volatile double a = (double)best[0] * w2;
            sp = sp + a;
            volatile double c = (double)best[1] * p4;
            sp = sp + c;
        }
    }
    return empty > 0 ? (float)(sp / (double)empty) : 0.0f;
}

ORC_API void orc_solve_layer(const uint64_t *keys, size_t n, float *rates_out,
                             const uint64_t *k1, const float *r1, size_t n1,
                             const uint64_t *k2, const float *r2, size_t n2,
                             const uint64_t *masks, int n_masks, const uint8_t *shifts, int n_shifts,
                             int symm, int target, int do_check, double p4, int fma_mode,
                             int nthreads) {
    orc_spec s;
    orc_book L1 = {(uint64_t *)k1, (float *)r1, n1}, L2 = {(uint64_t *)k2, (float *)r2, n2};
    orc_init();
    make_spec(&s, masks, n_masks, shifts, n_shifts, symm, target);
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 1024)
    for (size_t i = 0; i < n; ++i)
        rates_out[i] = solve_state(keys[i], &L1, &L2, &s, do_check, p4, fma_mode);
}

/* ------------------------------------------------------------------ */
/* Whole build, kept in memory                                          */
/* ------------------------------------------------------------------ */

typedef struct {
    int steps;
    uint64_t **raw;     /* raw layers 0..steps-1 (steps-1 is the partial layer) */
    size_t *raw_n;
    uint64_t *c2, *c4;  /* candidate counts per forward step (after pattern filter) */
    orc_book *book;     /* solved, zero-compacted layers 0..steps-1 */
} orc_build_t;

ORC_API void orc_free(orc_build_t *h) {
    if (!h) return;
    for (int i = 0; i < h->steps; ++i) {
        free(h->raw[i]);
        free(h->book[i].keys);
        free(h->book[i].rates);
    }
    free(h->raw);
    free(h->raw_n);
    free(h->c2);
    free(h->c4);
    free(h->book);
    free(h);
}

/* generate_process (BookGenerator.cpp:963-1121) as a set computation:
 * layer 0 = seeds in given order; for i = 1..steps-2, layer i = sorted
 * unique union of the spawn-2 children of layer i-1 and the spawn-4
 * children of layer i-2, parents of step i skipped iff i > docheck and
 * success.  The lossy hash prefilter is result-neutral and omitted; the
 * pivot-partitioned merge is a sorted union.  Layer steps-1 holds only the
 * spawn-4 children of layer steps-3.
 * Then final_steps (BookSolver.cpp:237-284) and recalculate_process_impl
 * (BookSolver.cpp:513-605). solve = 0 stops after the forward pass. */
ORC_API orc_build_t *orc_build(const uint64_t *seeds, size_t n_seeds, const uint64_t *masks,
                               int n_masks, const uint8_t *shifts, int n_shifts, int symm,
                               int target, int steps, int docheck, double p4, int fma_mode,
                               int solve, int nthreads) {
    orc_init();
    if (nthreads <= 0) nthreads = 1;
    if (steps < 3) return NULL;
    orc_spec s;
    make_spec(&s, masks, n_masks, shifts, n_shifts, symm, target);
    orc_build_t *h = (orc_build_t *)calloc(1, sizeof(orc_build_t));
    h->steps = steps;
    h->raw = (uint64_t **)calloc((size_t)steps, sizeof(uint64_t *));
    h->raw_n = (size_t *)calloc((size_t)steps, sizeof(size_t));
    h->c2 = (uint64_t *)calloc((size_t)steps, sizeof(uint64_t));
    h->c4 = (uint64_t *)calloc((size_t)steps, sizeof(uint64_t));
    h->book = (orc_book *)calloc((size_t)steps, sizeof(orc_book));

    h->raw[0] = (uint64_t *)malloc((n_seeds ? n_seeds : 1) * sizeof(uint64_t));
    memcpy(h->raw[0], seeds, n_seeds * sizeof(uint64_t));
    h->raw_n[0] = n_seeds;

    vec64 partial = {0, 0, 0}; /* spawn-4 candidates for layer i */
    for (int i = 1; i <= steps - 2; ++i) {
        vec64 c2 = {0, 0, 0}, c4 = {0, 0, 0};
        expand(h->raw[i - 1], h->raw_n[i - 1], &s, i > docheck, &c2, &c4, nthreads);
        h->c2[i] = c2.n;
        h->c4[i] = c4.n;
        for (size_t j = 0; j < partial.n; ++j) v_push(&c2, partial.p[j]);
        free(partial.p);
        size_t u = orc_sort_unique(c2.p, c2.n, nthreads);
        h->raw[i] = c2.p ? c2.p : (uint64_t *)malloc(8);
        h->raw_n[i] = u;
        partial = c4;
    }
    {
        size_t u = orc_sort_unique(partial.p, partial.n, nthreads);
        h->raw[steps - 1] = partial.p ? partial.p : (uint64_t *)malloc(8);
        h->raw_n[steps - 1] = u;
    }
    if (!solve) return h;

    /* final layers: success boards at rate 1.0, others dropped */
    for (int i = steps - 2; i <= steps - 1; ++i) {
        orc_book *B = &h->book[i];
        B->keys = (uint64_t *)malloc((h->raw_n[i] + 1) * sizeof(uint64_t));
        B->rates = (float *)malloc((h->raw_n[i] + 1) * sizeof(float));
        size_t m = 0;
        for (size_t j = 0; j < h->raw_n[i]; ++j)
            if (is_success(h->raw[i][j], &s)) {
                B->keys[m] = h->raw[i][j];
                B->rates[m] = 1.0f;
                ++m;
            }
        B->n = m;
    }
    for (int i = steps - 3; i >= 0; --i) {
        size_t n = h->raw_n[i];
        float *r = (float *)malloc((n + 1) * sizeof(float));
        orc_book *L1 = &h->book[i + 1], *L2 = &h->book[i + 2];
        int dc = i > docheck;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1024)
        for (size_t j = 0; j < n; ++j)
            r[j] = solve_state(h->raw[i][j], L1, L2, &s, dc, p4, fma_mode);
        /* compact_live_entries (BookSolver.cpp:380-394): keep rate > 0 in order */
        orc_book *B = &h->book[i];
        B->keys = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
        B->rates = (float *)malloc((n + 1) * sizeof(float));
        size_t m = 0;
        for (size_t j = 0; j < n; ++j)
            if (r[j] > 0.0f) {
                B->keys[m] = h->raw[i][j];
                B->rates[m] = r[j];
                ++m;
            }
        B->n = m;
        free(r);
    }
    return h;
}

ORC_API size_t orc_raw_layer(const orc_build_t *h, int i, const uint64_t **keys) {
    if (!h || i < 0 || i >= h->steps) { *keys = NULL; return 0; }
    *keys = h->raw[i];
    return h->raw_n[i];
}

ORC_API size_t orc_book_layer(const orc_build_t *h, int i, const uint64_t **keys, const float **rates) {
    if (!h || i < 0 || i >= h->steps) { *keys = NULL; *rates = NULL; return 0; }
    *keys = h->book[i].keys;
    *rates = h->book[i].rates;
    return h->book[i].n;
}

ORC_API void orc_candidates(const orc_build_t *h, int i, uint64_t *c2, uint64_t *c4) {
    *c2 = (h && i > 0 && i < h->steps) ? h->c2[i] : 0;
    *c4 = (h && i > 0 && i < h->steps) ? h->c4[i] : 0;
}
