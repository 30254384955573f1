/*
 * ad_oracle.c -- CPU restatement of the reference's advanced ("AD",
 * large-tile masking) forward pass: the masked mover, the masker's tile
 * statistics, derive / validate, and the layer-by-layer generation.
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
 *     (oracle/ref, tests/golden/ad_*.json).
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
