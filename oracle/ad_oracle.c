/*
 * ad_oracle.c -- CPU restatement of the reference's advanced ("AD",
 * large-tile masking) builder: the masked mover, the masker's tile
 * statistics, derive / validate and the layer-by-layer generation of the
 * forward pass, and the solver (bucketed rate rows over the placements of
 * the large tiles, float32).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle of the AD
 * forward pass: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.  The product path
 * never calls it.
 *
 * Pinning (how a plausible mistake is caught), in tests/:
 *   - the masked row table: exhaustive 65536-row check against an
 *     independent pure-Python restatement of merge_line;
 *   - the large-tile combination table: every entry against a brute-force
 *     search over multisets of large tiles;
 *   - whole forward passes: per-layer SHA-256 digests of the raw layer files
 *     written by the reference's own AD builder compiled in this container
 *     (oracle/ref, tests/golden/ad_*.json);
 *   - whole solves: per-bucket SHA-256 digests of the reference solver's
 *     <i>b/<key>.i|.b files (tests/golden/ad_book_*.json).
 *
 * Board encoding as in the classic path: nibble k = cell (row 3-k/4,
 * column 3-k%4), value v = tile 2^v, 0 empty, 0xF = a masked large tile.
 *
 * Set semantics.  The reference filters candidates through a lossy hash
 * (BookGeneratorAD.cpp:749-784, 640-703) and keeps every accepted board;
 * this restatement keeps the union of every candidate's boards (a plain
 * child itself, a child that made a new masked tile through derive), which
 * is what the reference's sorted, deduplicated layers hold as long as the
 * same board never arrives both with and without a new masked tile.  The
 * golden digests check exactly that.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* Masked mover (BoardMoverAD.cpp:47-86)                                */
/* ------------------------------------------------------------------ */

/* Slide a row toward its high nibble (cell 0 of the line is the high
 * nibble, CommonMover.h:52-58).  Equal neighbours merge once, a masked tile
 * (0xF) never merges, and a merge that reaches 64 or more becomes a masked
 * tile. */
static uint16_t ad_slide_row(uint16_t row) {
    int v[4], nz[4], n = 0, out[4] = {0, 0, 0, 0}, m = 0;
    for (int i = 0; i < 4; ++i) v[i] = (row >> (4 * (3 - i))) & 0xF;
    for (int i = 0; i < 4; ++i)
        if (v[i]) nz[n++] = v[i];
    for (int i = 0; i < n; ++i) {
        if (i + 1 < n && nz[i] == nz[i + 1] && nz[i] != 0xF) {
            int t = nz[i] + 1;
            out[m++] = t >= 6 ? 0xF : t;
            ++i;
        } else {
            out[m++] = nz[i];
        }
    }
    uint16_t r = 0;
    for (int i = 0; i < 4; ++i) r |= (uint16_t)(out[i] << (4 * (3 - i)));
    return r;
}

static uint16_t g_left[65536];
static int g_ready = 0;

static void ad_init(void) {
    if (g_ready) return;
    for (uint32_t r = 0; r < 65536; ++r) g_left[r] = ad_slide_row((uint16_t)r);
    g_ready = 1;
}

static inline uint64_t transpose(uint64_t x) {
    uint64_t t = ((x >> 24) ^ x) & 0x00000000FF00FF00ull;
    x ^= t ^ (t << 24);
    t = ((x >> 12) ^ x) & 0x0000F0F00000F0F0ull;
    x ^= t ^ (t << 12);
    return x;
}
static inline uint64_t mirror_lr(uint64_t x) {
    x = ((x >> 8) & 0x00FF00FF00FF00FFull) | ((x & 0x00FF00FF00FF00FFull) << 8);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
    return x;
}
static inline uint64_t mirror_ud(uint64_t x) {
    x = (x >> 32) | (x << 32);
    return ((x >> 16) & 0x0000FFFF0000FFFFull) | ((x & 0x0000FFFF0000FFFFull) << 16);
}
static inline uint64_t rows_left(uint64_t b) {
    return (uint64_t)g_left[b & 0xFFFF] | ((uint64_t)g_left[(b >> 16) & 0xFFFF] << 16) |
           ((uint64_t)g_left[(b >> 32) & 0xFFFF] << 32) | ((uint64_t)g_left[b >> 48] << 48);
}
/* number of masked (0xF) cells */
static inline int f_count(uint64_t b) {
    return __builtin_popcountll(b & (b >> 1) & (b >> 2) & (b >> 3) & 0x1111111111111111ull);
}

/* Moves in the reference's order left, right, up, down
 * (BookGeneratorAD.cpp:870-877).  A merge into a masked tile is the only
 * way the count of 0xF cells grows (a masked tile never merges), so a move
 * made a new masked tile exactly when the moved board has more of them;
 * this is the flag m_move_left2 / m_move_up2 read from their row bitmap
 * (BoardMoverAD.cpp:100-117,129-133). */
static inline void ad_move4(uint64_t b, uint64_t out[4], int mnt[4]) {
    const uint64_t t = transpose(b);
    const int f0 = f_count(b);
    out[0] = rows_left(b);
    out[1] = mirror_lr(rows_left(mirror_lr(b)));
    out[2] = transpose(rows_left(t));
    out[3] = transpose(mirror_lr(rows_left(mirror_lr(t))));
    for (int d = 0; d < 4; ++d) mnt[d] = f_count(out[d]) > f0;
}

static inline uint64_t umin(uint64_t a, uint64_t b) { return a < b ? a : b; }

/* apply_canonical (BookGeneratorAD.cpp:85-103), modes 0-3 */
static inline uint64_t canonical(uint64_t b, int mode) {
    if (mode == 0) return b;
    if (mode == 3) return umin(b, mirror_lr(b));
    const uint64_t t = transpose(b);
    if (mode == 2) return umin(b, t);
    uint64_t m = umin(b, t);
    const uint64_t lb = mirror_lr(b), ub = mirror_ud(b), lt = mirror_lr(t), ut = mirror_ud(t);
    m = umin(m, umin(lb, ub));
    m = umin(m, umin(mirror_ud(lb), lt));
    m = umin(m, umin(ut, mirror_ud(lt)));
    return m;
}

/* ------------------------------------------------------------------ */
/* Masker (BoardMaskerAD.cpp)                                           */
/* ------------------------------------------------------------------ */

typedef struct {
    uint64_t masks[16];
    int n_masks;
    int symm;
    uint64_t fixed_mask;   /* 0xF at every pinned 32768 cell (fixed_32k_mask, :96-102) */
    int n_fixed, n_free;   /* pinned and free 32768 tiles */
    uint32_t limit;        /* small_tile_sum_limit */
    int target;            /* log2 of the target tile */
} ad_spec;

/* masked_tiles_combinations (BoardMaskerAD.cpp:256-285) for
 * (remaining_sum = a << 6, remaining_count = b), stored ascending as
 * init_masker reverses it (:517-526).  Returns the length (0 = none). */
static uint8_t g_comb[2560][10];
static uint8_t g_comb_n[2560];

static int comb_build(uint32_t sum, int count, uint8_t *out) {
    static const uint32_t tiles[8] = {8192, 4096, 2048, 1024, 512, 256, 128, 64};
    uint8_t tmp[16];
    int n = 0;
    for (int k = 0; k < 8; ++k) {
        const uint32_t t = tiles[k];
        const uint8_t lg = (uint8_t)(13 - k);
        if (t > sum) continue;
        if (t == sum) {
            if (count == 1) {
                tmp[n++] = lg;
                goto done;
            }
        } else {
            if (2 * t == sum && count == 2) {
                tmp[n++] = lg;
                tmp[n++] = lg;
                goto done;
            }
            if (t != 64 && sum == 192 && count == 3) {
                tmp[n++] = 6;
                tmp[n++] = 6;
                tmp[n++] = 6;
                goto done;
            }
            sum -= t;
            --count;
            tmp[n++] = lg;
        }
    }
    return 0;
done:
    for (int i = 0; i < n; ++i) out[i] = tmp[n - 1 - i];
    return n;
}

static int g_comb_ready = 0;
static void comb_init(void) {
    if (g_comb_ready) return;
    memset(g_comb_n, 0, sizeof g_comb_n);
    for (int a = 1; a < 256; ++a)
        for (int b = 1; b < 10; ++b) g_comb_n[a + (b << 8)] = (uint8_t)comb_build((uint32_t)a << 6, b, g_comb[a + (b << 8)]);
    g_comb_ready = 1;
}

/* tiles_combination_view (BoardMaskerAD.h:157-160) with the reference's
 * uint8 truncation of both coordinates.  An index past the table (a count
 * byte of 10 or more) reads nothing here. */
static inline const uint8_t *comb_view(uint32_t large_sum, int count_minus_free, int *n) {
    const uint32_t idx = (uint32_t)(uint8_t)(large_sum >> 6) + ((uint32_t)(uint8_t)count_minus_free << 8);
    if (idx >= 2560) {
        *n = 0;
        return NULL;
    }
    *n = g_comb_n[idx];
    return g_comb[idx];
}

/* validate (BoardMaskerAD.cpp:422-456) with tile_sum_and_32k_count2's
 * statistics (:165-187): pinned cells cleared, cells of 64 or more (value
 * > 5) counted, the others summed. */
static int ad_validate(uint64_t b, uint32_t board_sum, const ad_spec *s) {
    b &= ~s->fixed_mask;
    uint32_t total = 0;
    int count = 0;
    for (int k = 0; k < 16; ++k) {
        const int v = (int)((b >> (4 * k)) & 0xF);
        if (v > 5) ++count;
        else if (v > 0) total += 1u << v;
    }
    if (total >= s->limit + 64u) return 0;
    if (count == s->n_free) return 1;
    const uint32_t large = board_sum - total - ((uint32_t)(s->n_free + s->n_fixed) << 15);
    int n;
    const uint8_t *c = comb_view(large, count - s->n_free, &n);
    if (n == 0) return 0;
    if (c[n - 1] == s->target) return 0;
    if (count < 2) return 1;
    if (n > 2 && c[0] == c[2]) {
        if (total >= s->limit - 64u) return 0;
    } else if (n > 1 && c[0] == c[1]) {
        if (total >= s->limit) return 0;
    }
    return 1;
}

/* derive (BookGeneratorAD.cpp:435-509) with tile_sum_and_32k_count3's
 * statistics (BoardMaskerAD.cpp:190-216).  Returns -1 for an invalid board,
 * 0 for a valid board kept as it is, else 1 with *m boards in out. */
static int ad_derive(uint64_t b, uint32_t board_sum, const ad_spec *s, uint64_t *out, int *m) {
    uint32_t total = 0;
    int count = 0, n64 = 0;
    int pos[16];
    uint64_t masked = b, pos64 = 0;
    for (int i = 60; i >= 0; i -= 4) {
        const int v = (int)((b >> i) & 0xF);
        if (v > 5) {
            if (((s->fixed_mask >> i) & 0xF) == 0) {
                pos[count++] = i;
                if (v == 6) {
                    ++n64;
                    pos64 = (uint64_t)i;
                }
            }
        } else if (v > 0) {
            total += 1u << v;
        }
    }
    if (n64 == 1) masked |= 0xFull << pos64;
    *m = 0;
    if (total >= s->limit + 64u) return -1;
    const uint32_t large = board_sum - total - ((uint32_t)(s->n_free + s->n_fixed) << 15);
    int n;
    const uint8_t *c = comb_view(large, count - s->n_free, &n);
    if (n == 0) return -1;
    if (c[n - 1] == s->target) return -1;
    if (count - s->n_free < 2) return 0;
    if (c[0] == c[1]) {
        if (total >= s->limit) return -1;
        if (n > 2 && c[0] == c[2]) {
            if (total >= s->limit - 64u) return -1;
            if (n64 == 0) return 1;  /* derived, no boards */
            /* derive_3x64 (:421-433): each large cell not already 64 set to 64 */
            for (int k = 0; k < count; ++k) {
                const uint64_t v = (b >> pos[k]) & 0xF;
                if (v == 6) continue;
                out[(*m)++] = b ^ ((v ^ 6ull) << pos[k]);
            }
            return 1;
        }
        /* two equal large tiles: every pair of large cells set to that tile */
        uint64_t delta[16];
        for (int k = 0; k < count; ++k) delta[k] = (((b >> pos[k]) & 0xF) ^ (uint64_t)c[0]) << pos[k];
        for (int p = 0; p < count - 1; ++p)
            for (int q = p + 1; q < count; ++q) out[(*m)++] = b ^ delta[p] ^ delta[q];
        return 1;
    }
    if (masked != b) {
        out[(*m)++] = masked;
        return 1;
    }
    return 0;
}

static inline int is_pattern(uint64_t b, const ad_spec *s) {
    if (s->n_masks == 0) return 1;
    for (int i = 0; i < s->n_masks; ++i)
        if ((b & s->masks[i]) == s->masks[i]) return 1;
    return 0;
}

/* mask_board (BoardMaskerAD.cpp:111-120): tiles of 64 or more to 0xF */
static inline uint64_t mask_board(uint64_t b) {
    uint64_t r = b;
    for (int k = 0; k < 16; ++k)
        if (((b >> (4 * k)) & 0xF) >= 6) r |= 0xFull << (4 * k);
    return r;
}

/* ------------------------------------------------------------------ */
/* Layer generation (BookGeneratorAD.cpp:785-960, 1256-1396)            */
/* ------------------------------------------------------------------ */

typedef struct {
    uint64_t *a;
    size_t n, cap;
} vec64;

static void v_push(vec64 *v, uint64_t x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 1024;
        v->a = (uint64_t *)realloc(v->a, v->cap * 8);
        if (!v->a) abort();
    }
    v->a[v->n++] = x;
}

static int cmp_u64(const void *x, const void *y) {
    const uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
    return a < b ? -1 : a > b;
}
static void sort_unique(vec64 *v) {
    if (v->n < 2) return;
    qsort(v->a, v->n, 8, cmp_u64);
    size_t m = 1;
    for (size_t i = 1; i < v->n; ++i)
        if (v->a[i] != v->a[m - 1]) v->a[m++] = v->a[i];
    v->n = m;
}

/* One candidate's boards (process_buffered_accepted_boards, :640-703) */
static void emit(uint64_t canon, int derive_kind, uint32_t child_sum, const ad_spec *s, vec64 *out) {
    if (!derive_kind) {
        v_push(out, canon);
        return;
    }
    uint64_t d[128];
    int m;
    const int r = ad_derive(canon, child_sum, s, d, &m);
    if (r < 0) return;
    if (r == 0) {
        v_push(out, canon);
        return;
    }
    for (int k = 0; k < m; ++k)
        if (is_pattern(d[k], s)) v_push(out, canonical(d[k], s->symm));
}

/* gen_boards_ad: children of every parent, spawn 2 into out2 and spawn 4
 * into out4, in the reference's cell / spawn / move order */
static void gen_layer(const uint64_t *par, size_t n, uint32_t board_sum, const ad_spec *s, vec64 *out2, vec64 *out4) {
    for (size_t p = 0; p < n; ++p) {
        const uint64_t t = par[p];
        for (int i = 0; i < 16; ++i) {
            if ((t >> (4 * i)) & 0xF) continue;
            for (int sp = 1; sp <= 2; ++sp) {
                const uint64_t t1 = t | ((uint64_t)sp << (4 * i));
                uint64_t mv[4];
                int mnt[4];
                ad_move4(t1, mv, mnt);
                for (int d = 0; d < 4; ++d) {
                    if (mv[d] == t1 || !is_pattern(mv[d], s)) continue;
                    emit(canonical(mv[d], s->symm), mnt[d], board_sum + (sp == 1 ? 2u : 4u), s, sp == 1 ? out2 : out4);
                }
            }
        }
    }
}

static void validate_vec(vec64 *v, uint32_t board_sum, const ad_spec *s) {
    size_t m = 0;
    for (size_t i = 0; i < v->n; ++i)
        if (ad_validate(v->a[i], board_sum, s)) v->a[m++] = v->a[i];
    v->n = m;
}

static int write_raw(const char *path, const uint64_t *a, size_t n) {
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    if (n && fwrite(a, 8, n, f) != n) {
        fclose(f);
        return -1;
    }
    return fclose(f);
}

static void make_spec(ad_spec *s, const uint64_t *masks, int n_masks, int symm, int n_free, const uint8_t *fixed_shifts,
                      int n_fixed, int limit, int target) {
    memset(s, 0, sizeof *s);
    s->n_masks = n_masks > 16 ? 16 : n_masks;
    for (int i = 0; i < s->n_masks; ++i) s->masks[i] = masks[i];
    s->symm = symm;
    s->n_free = n_free;
    s->n_fixed = n_fixed;
    for (int i = 0; i < n_fixed; ++i) s->fixed_mask += 0xFull << fixed_shifts[i];
    s->limit = (uint32_t)limit;
    s->target = target;
}

/* generate_process_ad (BookGeneratorAD.cpp:1256-1396) for a fresh build:
 * writes raw layers <pathname>0 .. <pathname>steps-2 (layer 0 = the masked
 * seeds as given, layer i = sorted unique union of layer i-2's spawn-4 and
 * layer i-1's spawn-2 children), validating the two layers at hand every
 * 32 steps as the reference does.  Returns 0, or -1 on an I/O error. */
ORC_API int orc_ad_forward(const uint64_t *seeds, size_t n_seeds, const uint64_t *masks, int n_masks, int symm,
                           int n_free, const uint8_t *fixed_shifts, int n_fixed, int limit, int target, int steps,
                           const char *pathname, uint64_t *sizes) {
    ad_init();
    comb_init();
    ad_spec s;
    make_spec(&s, masks, n_masks, symm, n_free, fixed_shifts, n_fixed, limit, target);
    uint32_t ini_sum = 0;
    for (int k = 0; k < 16 && n_seeds; ++k) {
        const uint32_t v = (uint32_t)((seeds[0] >> (4 * k)) & 0xF);
        if (v) ini_sum += 1u << v;
    }
    vec64 d0 = {0}, d1 = {0};
    for (size_t i = 0; i < n_seeds; ++i) v_push(&d0, mask_board(seeds[i]));
    char path[4096];
    snprintf(path, sizeof path, "%s0", pathname);
    if (write_raw(path, d0.a, d0.n)) return -1;
    if (sizes) sizes[0] = d0.n;
    for (int i = 1; i < steps - 1; ++i) {
        const uint32_t board_sum = (uint32_t)(2 * i) + ini_sum - 2u;
        vec64 a1 = {0}, a2 = {0};
        gen_layer(d0.a, d0.n, board_sum, &s, &a1, &a2);
        sort_unique(&a1);
        sort_unique(&a2);
        /* layer i = d1 (spawn-4 children of step i-1) merged with a1 */
        for (size_t k = 0; k < d1.n; ++k) v_push(&a1, d1.a[k]);
        sort_unique(&a1);
        free(d0.a);
        free(d1.a);
        d0 = a1;
        d1 = a2;
        if (((i + (int)(ini_sum % 64u / 2u)) % 32) == (int)((s.limit / 2u) % 32u) + 1) {
            validate_vec(&d0, board_sum + 2u, &s);
            validate_vec(&d1, board_sum + 4u, &s);
        }
        snprintf(path, sizeof path, "%s%d", pathname, i);
        if (write_raw(path, d0.a, d0.n)) return -1;
        if (sizes) sizes[i] = d0.n;
    }
    free(d0.a);
    free(d1.a);
    return 0;
}

/* Primitives for the tests */
ORC_API uint16_t orc_ad_row_left(uint16_t row) { return ad_slide_row(row); }

ORC_API void orc_ad_move4(uint64_t b, uint64_t *out, int32_t *mnt) {
    ad_init();
    int m[4];
    ad_move4(b, out, m);
    for (int d = 0; d < 4; ++d) mnt[d] = m[d];
}

/* the combination table as [2560][11]: length, then up to 10 tiles */
ORC_API void orc_ad_comb_table(uint8_t *out) {
    comb_init();
    for (int i = 0; i < 2560; ++i) {
        out[i * 11] = g_comb_n[i];
        memcpy(out + i * 11 + 1, g_comb[i], 10);
    }
}

ORC_API int orc_ad_derive(uint64_t b, uint32_t board_sum, const uint64_t *masks, int n_masks, int symm, int n_free,
                          const uint8_t *fixed_shifts, int n_fixed, int limit, int target, uint64_t *out, int32_t *m) {
    comb_init();
    ad_spec s;
    make_spec(&s, masks, n_masks, symm, n_free, fixed_shifts, n_fixed, limit, target);
    int mm = 0;
    const int r = ad_derive(b, board_sum, &s, out, &mm);
    *m = mm;
    return r;
}

ORC_API int orc_ad_validate(uint64_t b, uint32_t board_sum, const uint64_t *masks, int n_masks, int symm, int n_free,
                            const uint8_t *fixed_shifts, int n_fixed, int limit, int target) {
    comb_init();
    ad_spec s;
    make_spec(&s, masks, n_masks, symm, n_free, fixed_shifts, n_fixed, limit, target);
    return ad_validate(b, board_sum, &s);
}

/* ------------------------------------------------------------------ */
/* Solver (BookSolverAD.cpp): float32, float64 or uint32 rates         */
/* ------------------------------------------------------------------ */
/*
 * recalculate_process_ad_impl (:2086-2242) over the raw layers the forward
 * pass above wrote, layers steps-3 .. 0, each written as bucket files
 * <pathname><i>b/<key>.i / <key>.b (dict_tofile, :320-352).  Pinned by the
 * per-bucket digests of the reference's own solver (tests/golden/
 * ad_book_*.json).  Serial, one board at a time, the placements' column
 * order computed for each move (the reference memoises it).  Rates are held
 * as doubles, which hold every float32 and uint32 value exactly, and take
 * the rate type's rounding where the reference converts (one conversion per
 * stored rate).
 */
#include <math.h>
#include <sys/stat.h>

#define KEY_MIN (-16)
#define N_KEYS 48

typedef struct {
    uint64_t *ind;
    double *rates;
    size_t rows;
    uint32_t cols;
} sbucket;

/* the rate type: 0 float32, 1 float64, 2 uint32 (scale 4e9) */
static int g_rt = 0;
static double rt_max(void) { return g_rt == 2 ? 4000000000.0 : 1.0; }
static double rt_store(double x) {  /* static_cast<T>(double) */
    if (g_rt == 0) return (double)(float)x;
    if (g_rt == 2) return (double)(uint32_t)x;
    return x;
}
typedef struct {
    sbucket b[N_KEYS];
} slayer;

static uint16_t g_classic[65536];
static int g_classic_ready = 0;

/* StandardMergePolicy (BoardMover.h:7-55): compact, equal neighbours merge
 * once from the move side, 32768 (0xF) never merges */
static void classic_init(void) {
    if (g_classic_ready) return;
    for (uint32_t row = 0; row < 65536; ++row) {
        int t[4], n = 0, out[4] = {0, 0, 0, 0}, m = 0;
        for (int i = 0; i < 4; ++i) {
            const int v = (int)((row >> (4 * (3 - i))) & 0xF);
            if (v) t[n++] = v;
        }
        for (int i = 0; i < n; ++i) {
            if (i + 1 < n && t[i] == t[i + 1] && t[i] != 15) {
                out[m++] = t[i] + 1;
                ++i;
            } else {
                out[m++] = t[i];
            }
        }
        g_classic[row] = (uint16_t)((out[0] << 12) | (out[1] << 8) | (out[2] << 4) | out[3]);
    }
    g_classic_ready = 1;
}

static uint64_t classic_rows_left(uint64_t b) {
    uint64_t r = 0;
    for (int i = 0; i < 4; ++i) r |= (uint64_t)g_classic[(b >> (16 * i)) & 0xFFFF] << (16 * i);
    return r;
}

/* move_board(b, d + 1): 0 left, 1 right, 2 up, 3 down (CommonMover.h:126-134) */
static uint64_t classic_move(uint64_t b, int d) {
    switch (d) {
        case 0: return classic_rows_left(b);
        case 1: return mirror_lr(classic_rows_left(mirror_lr(b)));
        case 2: return transpose(classic_rows_left(transpose(b)));
        default: return transpose(mirror_lr(classic_rows_left(mirror_lr(transpose(b)))));
    }
}

/* apply_sym_like (SymmetryUtils.h:9-29) as cell maps: cell (r, c) = nibble
 * 4r + c goes to the cell below */
static uint64_t sym_like(uint64_t b, int idx) {
    if (idx == 0) return b;
    uint64_t out = 0;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            int nr = r, nc = c;
            switch (idx) {
                case 1: nc = 3 - c; break;                 /* ReverseLR */
                case 2: nr = 3 - r; break;                 /* ReverseUD */
                case 3: nr = c; nc = r; break;             /* ReverseUL */
                case 4: nr = 3 - c; nc = 3 - r; break;     /* ReverseUR */
                case 5: nr = 3 - r; nc = 3 - c; break;     /* Rotate180 */
                case 6: nr = 3 - c; nc = r; break;         /* RotateL */
                default: nr = c; nc = 3 - r; break;        /* RotateR */
            }
            out |= ((b >> (4 * (4 * r + c))) & 0xF) << (4 * (4 * nr + nc));
        }
    return out;
}

/* canonical_*_pair (Calculator.h:103-128): the first strictly smaller */
static uint64_t sym_pair(uint64_t b, int mode, int *idx) {
    static const int full[7] = {1, 2, 3, 4, 5, 6, 7};
    *idx = 0;
    if (mode == 0) return b;
    const int *cand = mode == 2 ? (const int[]){3} : mode == 3 ? (const int[]){1} : full;
    const int nc = mode == 1 ? 7 : 1;
    uint64_t m = b;
    for (int i = 0; i < nc; ++i) {
        const uint64_t t = sym_like(b, cand[i]);
        if (t < m) {
            m = t;
            *idx = cand[i];
        }
    }
    return m;
}

/* the info tables' statistics (BoardMaskerAD.cpp:24-50,169-257) */
typedef struct {
    uint32_t total;
    int count, pos_rank, merged_found;
    uint64_t pos_bitmap;
    uint32_t merged_tile;
} tstats;

static tstats stats_large(uint64_t b, const ad_spec *s) {  /* count2 / count4: large = above 32 */
    b &= ~s->fixed_mask;
    tstats t = {0, 0, 0, 0, 0, 0};
    int seen = 0;
    for (int k = 15; k >= 0; --k) {
        const uint32_t v = (uint32_t)((b >> (4 * k)) & 0xF);
        if (v > 5) {
            ++t.count;
            t.pos_bitmap |= 0xFull << (4 * k);
            if (v == 15) {
                if (!seen) ++t.pos_rank;
            } else {
                ++t.merged_found;
                t.merged_tile ^= v;
                seen = 1;
            }
        } else if (v) {
            t.total += 1u << v;
        }
    }
    return t;
}

/* tile_sum_and_32k_count_positions: 0xF cells (highest first) and the sum of
 * the others, 16 bits a row */
static int f_positions(uint64_t b, const ad_spec *s, int pos[16], uint32_t *total) {
    b &= ~s->fixed_mask;
    int c = 0;
    *total = 0;
    for (int r = 3; r >= 0; --r) {
        uint32_t rs = 0;
        for (int i = 3; i >= 0; --i) {
            const int k = 4 * r + i;
            const uint32_t v = (uint32_t)((b >> (4 * k)) & 0xF);
            if (v == 15) pos[c++] = 4 * k;
            else if (v) rs += 1u << v;
        }
        *total += rs & 0xFFFF;
    }
    return c;
}

/* generate_permutations + resort_permutations + build_permutation_subsets */
typedef struct {
    uint8_t *rows;
    int n_rows, cols;
} pslot;
static pslot g_perm[256];

static void perm_dfs(int m, int n, uint8_t *cur, int depth, int *used, uint8_t *buf, int *cnt) {
    if (depth == n) {
        memcpy(buf + (size_t)(*cnt) * n, cur, (size_t)n);
        ++*cnt;
        return;
    }
    for (int i = 0; i < m; ++i) {
        if (used[i]) continue;
        used[i] = 1;
        cur[depth] = (uint8_t)i;
        perm_dfs(m, n, cur, depth + 1, used, buf, cnt);
        used[i] = 0;
    }
}

static int g_perm_m;
static uint64_t perm_key(const uint8_t *row, int m, int n) {
    uint64_t clear = 0, set = 0;
    for (int j = 0; j < n; ++j) {
        const int sh = (m - 1 - row[j]) * 4;
        clear += 0xFull << sh;
        set += (uint64_t)j << sh;
    }
    return ~clear | set;
}

static void perms_init(int target, int n_free) {
    for (int i = 0; i < 256; ++i) {
        free(g_perm[i].rows);
        g_perm[i].rows = NULL;
        g_perm[i].n_rows = g_perm[i].cols = 0;
    }
    for (int n = 1; n < target - 4; ++n) {
        const int m = n + n_free;
        if (m > 15) break;
        long total = 1;
        for (int k = 0; k < n; ++k) total *= (m - k);
        if (total > (1 << 16)) break;
        uint8_t *buf = malloc((size_t)total * n), cur[16];
        int used[16] = {0}, cnt = 0;
        perm_dfs(m, n, cur, 0, used, buf, &cnt);
        /* sort rows by key (insertion into an index order; rows are unique) */
        uint64_t *keys = malloc(sizeof(uint64_t) * cnt);
        int *order = malloc(sizeof(int) * cnt);
        for (int i = 0; i < cnt; ++i) {
            keys[i] = perm_key(buf + (size_t)i * n, m, n);
            order[i] = i;
        }
        for (int i = 1; i < cnt; ++i) {
            const int o = order[i];
            int j = i - 1;
            while (j >= 0 && keys[order[j]] > keys[o]) {
                order[j + 1] = order[j];
                --j;
            }
            order[j + 1] = o;
        }
        uint8_t *rows = malloc((size_t)cnt * n);
        for (int i = 0; i < cnt; ++i) memcpy(rows + (size_t)i * n, buf + (size_t)order[i] * n, (size_t)n);
        g_perm[(m << 4) + n] = (pslot){rows, cnt, n};
        free(buf);
        free(keys);
        free(order);
    }
    (void)g_perm_m;
}

/* unmask_board_into (BoardMaskerAD.cpp:445-500) */
static int unmask(uint64_t b, uint32_t sum, const ad_spec *s, uint64_t *out, int cap) {
    int pos[16];
    uint32_t total;
    const int cnt = f_positions(b, s, pos, &total);
    if (cnt == 0 || cnt == s->n_free) {
        out[0] = b;
        return 1;
    }
    int nt = 0;
    const uint8_t *tiles = comb_view(sum - total - ((uint32_t)(s->n_free + s->n_fixed) << 15), cnt - s->n_free, &nt);
    if (!tiles || nt == 0) return 0;
    if (cnt > 15 || cnt - s->n_free < 1) return 0;
    const pslot *p = &g_perm[(cnt << 4) + (cnt - s->n_free)];
    if (!p->n_rows) return 0;
    int m = 0;
    for (int r = 0; r < p->n_rows; ++r) {
        const uint8_t *row = p->rows + (size_t)r * p->cols;
        if (nt > 1 && tiles[0] == tiles[1] && !(row[1] < row[0])) continue;
        uint64_t clear = 0, set = 0;
        for (int j = 0; j < nt; ++j) {
            clear += 0xFull << pos[row[j]];
            set += (uint64_t)tiles[j] << pos[row[j]];
        }
        if (m < cap) out[m] = (b & ~clear) | set;
        ++m;
    }
    return m;
}

static void ranks(const uint64_t *b, int n, int *r) {
    for (int i = 0; i < n; ++i) {
        int k = 0;
        for (int j = 0; j < n; ++j) k += b[j] < b[i] || (b[j] == b[i] && j < i);
        r[i] = k;
    }
}

static const double *find_row(const slayer *L, int key, uint64_t board, uint32_t *cols) {
    if (key < KEY_MIN || key >= KEY_MIN + N_KEYS) return NULL;
    const sbucket *B = &L->b[key - KEY_MIN];
    size_t lo = 0, hi = B->rows;
    while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if (B->ind[mid] < board) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= B->rows || B->ind[lo] != board || !B->cols) return NULL;
    *cols = B->cols;
    return B->rates + lo * B->cols;
}

static uint32_t fact(int a) {
    static const uint32_t f[19] = {1, 1, 2, 6, 24, 120, 720, 5040, 40320, 362880, 3628800, 39916800, 1, 1, 1, 1, 1, 1, 1};
    return (a >= 0 && a < 19) ? f[a] : 0;
}

static int bucket_of(uint64_t b, uint32_t sum, const ad_spec *s) {
    const tstats t = stats_large(b, s);
    if (t.count == s->n_free) return t.count;
    int nt = 0;
    const uint8_t *c = comb_view(sum - t.total - ((uint32_t)(s->n_free + s->n_fixed) << 15), t.count - s->n_free, &nt);
    if (!c || !nt) return -1000;
    if (t.count - s->n_free == 1) return t.count;
    if (nt > 1 && c[0] == c[1]) return (nt > 2 && c[0] == c[2]) ? t.count - 3 + 16 : -t.count;
    return t.count;
}

static void find3x64(uint64_t un, const ad_spec *s, int *p64, int *p128, int *prank, int *pos64) {
    *p64 = *p128 = *prank = *pos64 = 0;
    for (int sh = 60; sh >= 0; sh -= 4) {
        const uint32_t v = (uint32_t)((un >> sh) & 0xF);
        if (v == 15) {
            if (!((s->fixed_mask >> sh) & 0xF)) ++*prank;
        } else if (v == 6) {
            *p64 = *prank;
            *pos64 = sh;
            ++*prank;
        } else if (v == 7) {
            *p128 = *prank;
            ++*prank;
        }
    }
}

static int map364(int x, int y, int n) {
    return x < y ? 2 * n * x - x * x - 2 * x + y - 1 : 2 * n * y - y * y - 3 * y + n + x - 2;
}

#define MAXD 64

/* solve_optimal_success_rate(_arr_into) for one spawned board: osr[0..d) */
static void solve_spawn(uint64_t b, uint64_t v, int pos, uint32_t sum, int key, int d, const slayer *L,
                        const ad_spec *s, double *osr) {
    const uint64_t bs = b | (v << (4 * pos));
    for (int j = 0; j < d; ++j) osr[j] = 0.0;
    uint64_t mv[4];
    int mnt[4];
    ad_move4(bs, mv, mnt);
    mnt[1] = mnt[0];  /* right takes left's flag, down up's (m_move_all_dir2, BoardMoverAD.cpp:180-186) */
    mnt[3] = mnt[2];
    for (int dir = 0; dir < 4; ++dir) {
        if (mv[dir] == bs || !is_pattern(mv[dir], s)) continue;
        int si;
        const uint64_t canon = sym_pair(mv[dir], s->symm, &si);
        uint64_t D[MAXD];
        int R[MAXD];
        int nd = unmask(bs, sum, s, D, MAXD);
        if (nd > MAXD) nd = MAXD;
        for (int j = 0; j < nd; ++j) D[j] = sym_like(classic_move(D[j], dir), si);
        ranks(D, nd, R);
        uint32_t cols = 0;
        if (!mnt[dir]) {
            const double *row = find_row(L, key, canon, &cols);
            if (!row) continue;
            if (d == 1) { osr[0] = osr[0] < row[0] ? row[0] : osr[0]; continue; }
            for (int i = 0; i < nd && i < d; ++i)
                if (row[R[i]] > osr[i]) osr[i] = row[R[i]];
            continue;
        }
        const uint64_t un = sym_like(classic_move(bs, dir), si);
        if (key > 15) {  /* update_mnt_osr_364_ad / _arr_ad */
            int p64, p128, prank, pos64;
            find3x64(un, s, &p64, &p128, &prank, &pos64);
            const double *row = find_row(L, prank, canon | (0xFull << pos64), &cols);
            if (!row) continue;
            if (d == 1) {
                const double x = row[map364(p64, p128, prank)];
                osr[0] = osr[0] < x ? x : osr[0];
                continue;
            }
            const pslot *p = &g_perm[(prank << 4) + (prank - s->n_free)];
            int sub[MAXD], ns = 0;
            for (int r = 0; r < p->n_rows && ns < MAXD; ++r)
                if (p->rows[(size_t)r * p->cols] == p64 && p->rows[(size_t)r * p->cols + 1] == p128) sub[ns++] = r;
            if (ns < nd) continue;
            for (int i = 0; i < nd && i < d; ++i)
                if (row[sub[R[i]]] > osr[i]) osr[i] = row[sub[R[i]]];
            continue;
        }
        const tstats t = stats_large(un, s);
        const int success = t.merged_tile == (uint32_t)s->target && t.merged_found == 1;
        if (success) {
            for (int j = 0; j < d; ++j) osr[j] = rt_max();
            continue;
        }
        int nt = 0;
        const uint8_t *c = comb_view(sum - t.total - ((uint32_t)(s->n_free + s->n_fixed) << 15), t.count - s->n_free, &nt);
        if (!c || !nt) continue;
        if (d == 1) {  /* update_mnt_osr_ad */
            if (t.merged_found == 2) continue;
            int bk = t.count;
            uint64_t board = canon;
            if (nt > 1 && c[0] == c[1]) {
                if (nt > 2 && c[0] == c[2]) {
                    bk = t.count + 13;
                    board = un;
                } else {
                    bk = -t.count;
                    board = (canon & ~t.pos_bitmap) | (((uint64_t)c[0] * 0x1111111111111111ull) & t.pos_bitmap);
                }
            }
            const double *row = find_row(L, bk, board, &cols);
            if (!row) continue;
            const double x = (bk > 0 && bk <= 15) ? row[t.pos_rank] : row[0];
            osr[0] = osr[0] < x ? x : osr[0];
            continue;
        }
        if (nt > 2 && c[0] == c[1] && c[0] == c[2]) {  /* arr3 */
            const double *row = find_row(L, t.count + 13, un, &cols);
            if (!row) continue;
            for (int i = 0; i < nd && i < d; ++i)
                if (row[R[i]] > osr[i]) osr[i] = row[R[i]];
        } else if (nt > 1 && c[0] == c[1]) {  /* arr1 */
            int lp[16], nl = 0;
            for (int k = 15; k >= 0; --k)
                if ((t.pos_bitmap >> (4 * k)) & 0xF) lp[nl++] = 4 * k;
            for (int j = 0; j < t.count && j < nl; ++j) {
                if (j == t.pos_rank) continue;
                const uint64_t ubj = (un & ~(0xFull << lp[j])) | ((uint64_t)c[0] << lp[j]);
                int si2;
                const uint64_t cb = sym_pair(ubj, s->symm, &si2);
                const double *row = find_row(L, -t.count, cb, &cols);
                if (!row) continue;
                uint64_t mb[MAXD];
                int mi[MAXD], nm = 0, R2[MAXD];
                for (int q = 0; q < nd; ++q)
                    if (((D[q] >> lp[j]) & 0xF) == t.merged_tile) {
                        mi[nm] = q;
                        mb[nm++] = sym_like(D[q], si2);
                    }
                ranks(mb, nm, R2);
                for (int q = 0; q < nm; ++q)
                    if (row[R2[q]] > osr[mi[q]]) osr[mi[q]] = row[R2[q]];
            }
        } else {  /* arr2 */
            const double *row = find_row(L, t.count, canon, &cols);
            if (!row) continue;
            const pslot *p = &g_perm[(t.count << 4) + (t.count - s->n_free)];
            int sub[MAXD], ns = 0;
            for (int r = 0; r < p->n_rows && ns < MAXD; ++r)
                if (p->rows[(size_t)r * p->cols] == t.pos_rank) sub[ns++] = r;
            if (ns < nd) continue;
            for (int i = 0; i < nd && i < d; ++i)
                if (row[sub[R[i]]] > osr[i]) osr[i] = row[sub[R[i]]];
        }
    }
}

static void layer_free(slayer *L) {
    for (int k = 0; k < N_KEYS; ++k) {
        free(L->b[k].ind);
        free(L->b[k].rates);
    }
    memset(L, 0, sizeof *L);
}

static int write_bucket_files(const char *folder, const slayer *L) {
    mkdir(folder, 0777);
    char path[4400];
    if (strlen(folder) > 4300) return -1;
    for (int k = 0; k < N_KEYS; ++k) {
        const sbucket *B = &L->b[k];
        if (!B->rows) continue;
        if (snprintf(path, sizeof path, "%.4300s/%d.i", folder, k + KEY_MIN) >= (int)sizeof path) return -1;
        if (write_raw(path, B->ind, B->rows)) return -1;
        if (snprintf(path, sizeof path, "%.4300s/%d.b", folder, k + KEY_MIN) >= (int)sizeof path) return -1;
        FILE *f = fopen(path, "wb");
        if (!f) return -1;
        const size_t m = B->rows * B->cols;
        for (size_t i = 0; i < m; ++i) {  /* the rate type's bytes, little-endian */
            if (g_rt == 0) {
                const float v = (float)B->rates[i];
                fwrite(&v, sizeof v, 1, f);
            } else if (g_rt == 2) {
                const uint32_t v = (uint32_t)B->rates[i];
                fwrite(&v, sizeof v, 1, f);
            } else {
                fwrite(&B->rates[i], sizeof(double), 1, f);
            }
        }
        fclose(f);
    }
    return 0;
}

/* The whole advanced build: the forward pass (orc_ad_forward) and the solve
 * of layers steps-3 .. 0 into <pathname><i>b, rate type rate_type (0
 * float32, 1 float64, 2 uint32).  Returns 0, or -1. */
ORC_API int orc_ad_build(const uint64_t *seeds, size_t n_seeds, const uint64_t *masks, int n_masks, int symm,
                         int n_free, const uint8_t *fixed_shifts, int n_fixed, int limit, int target, int steps,
                         const char *pathname, double p4, int rate_type) {
    if (rate_type < 0 || rate_type > 2) return -1;
    g_rt = rate_type;
    if (orc_ad_forward(seeds, n_seeds, masks, n_masks, symm, n_free, fixed_shifts, n_fixed, limit, target, steps,
                       pathname, NULL))
        return -1;
    classic_init();
    ad_spec s;
    make_spec(&s, masks, n_masks, symm, n_free, fixed_shifts, n_fixed, limit, target);
    perms_init(target, n_free);
    uint32_t ini = 0;
    for (int k = 0; k < 16 && n_seeds; ++k) {
        const uint32_t v = (uint32_t)((seeds[0] >> (4 * k)) & 0xF);
        if (v) ini += 1u << v;
    }
    const double w4 = p4, w2 = 1.0 - p4;
    slayer L1, L2;
    memset(&L1, 0, sizeof L1);
    memset(&L2, 0, sizeof L2);
    char path[4200];
    for (int step = steps - 3; step >= 0; --step) {
        snprintf(path, sizeof path, "%s%d", pathname, step);
        FILE *f = fopen(path, "rb");
        if (!f) return -1;
        fseek(f, 0, SEEK_END);
        const size_t n = (size_t)ftell(f) / 8;
        fseek(f, 0, SEEK_SET);
        uint64_t *raw = malloc(n * 8 + 8);
        if (fread(raw, 8, n, f) != n) {
            fclose(f);
            return -1;
        }
        fclose(f);
        const uint32_t sum = (uint32_t)(2 * step) + ini;
        slayer L0;
        memset(&L0, 0, sizeof L0);
        int *keys = malloc(sizeof(int) * (n + 1));
        for (size_t i = 0; i < n; ++i) {
            keys[i] = bucket_of(raw[i], sum, &s);
            if (keys[i] < KEY_MIN || keys[i] >= KEY_MIN + N_KEYS) keys[i] = -1000;
            else L0.b[keys[i] - KEY_MIN].rows++;
        }
        for (int k = 0; k < N_KEYS; ++k) {
            sbucket *B = &L0.b[k];
            const int key = k + KEY_MIN;
            const int a = key < 0 ? -key - 2 : key > 15 ? key - 16 : key;
            B->cols = fact(a) / fact(n_free);
            B->ind = malloc(B->rows * 8 + 8);
            B->rates = malloc(B->rows * B->cols * sizeof(double) + 8);
            B->rows = 0;
        }
        for (size_t i = 0; i < n; ++i)
            if (keys[i] != -1000) {
                sbucket *B = &L0.b[keys[i] - KEY_MIN];
                B->ind[B->rows++] = raw[i];
            }
        free(keys);
        free(raw);
        for (int k = 0; k < N_KEYS; ++k) {
            sbucket *B = &L0.b[k];
            const int key = k + KEY_MIN, d = (int)B->cols;
            if (!d || d > MAXD) continue;
            for (size_t r = 0; r < B->rows; ++r) {
                const uint64_t b = B->ind[r];
                double sp[MAXD];
                double o2[MAXD], o4[MAXD];
                int empty = 0;
                for (int j = 0; j < d; ++j) sp[j] = 0.0;
                for (int pos = 0; pos < 16; ++pos) {
                    if ((b >> (4 * pos)) & 0xF) continue;
                    ++empty;
                    solve_spawn(b, 1, pos, sum + 2u, key, d, &L1, &s, o2);
                    solve_spawn(b, 2, pos, sum + 4u, key, d, &L2, &s, o4);
                    if (d == 1) {  /* -Ofast: sum + fma(r2, 1 - p4, r4 * p4) */
                        sp[0] = sp[0] + fma(o2[0], w2, o4[0] * w4);
                    } else {       /* per spawn: fma(r, w, sum) */
                        for (int j = 0; j < d; ++j) sp[j] = fma(o2[j], w2, sp[j]);
                        for (int j = 0; j < d; ++j) sp[j] = fma(o4[j], w4, sp[j]);
                    }
                }
                double *dst = B->rates + r * d;
                if (d == 1) dst[0] = empty ? rt_store(sp[0] / (double)empty) : 0.0;
                else {
                    const double inv = empty ? 1.0 / (double)empty : 0.0;
                    for (int j = 0; j < d; ++j) dst[j] = empty ? rt_store(sp[j] * inv) : 0.0;
                }
            }
        }
        /* remove_died_ad: rows with no rate above zero */
        for (int k = 0; k < N_KEYS; ++k) {
            sbucket *B = &L0.b[k];
            if (!B->rows || !B->cols) continue;
            size_t m = 0;
            for (size_t r = 0; r < B->rows; ++r) {
                int keep = 0;
                for (uint32_t j = 0; j < B->cols; ++j) keep |= B->rates[r * B->cols + j] > 0.0;
                if (!keep) continue;
                B->ind[m] = B->ind[r];
                memmove(B->rates + m * B->cols, B->rates + r * B->cols, B->cols * sizeof(double));
                ++m;
            }
            B->rows = m;
            if (!m) B->cols = 0;
        }
        snprintf(path, sizeof path, "%s%db", pathname, step);
        if (write_bucket_files(path, &L0)) return -1;
        layer_free(&L2);
        L2 = L1;
        L1 = L0;
    }
    layer_free(&L1);
    layer_free(&L2);
    return 0;
}
