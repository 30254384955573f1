// Advanced ("AD") solver: success rates of the masked layers, steps-3 .. 0.
//
// The reference's advanced solver (native_core/src/BookSolverAD.cpp) keeps a
// layer as buckets keyed by its large tiles (expand_ad, :1243-1318): a board
// whose large tiles are all free 32768s, or all but one, or whose large tiles
// are distinct has key = their count; a board that holds a spelled-out equal
// pair has key -count, one with spelled-out 64s count + 13.  A board of a
// bucket stands for every placement of its large tiles over its 0xF cells, so
// its row holds one rate per placement: derive_size (:291-299) columns, in
// the order of the permutation table's rows (generate_permutations +
// resort_permutations, BoardMaskerAD.cpp:291-358).
//
// recalculate_ad (:1321-1505), per board: for every empty cell and spawn
// (2 with layer i+1, 4 with layer i+2), the best rate over the four masked
// moves.  A move that made no new masked tile looks its canonical board up
// in the same bucket; with the column order of the moved placements
// (their boards ranked, match_arr_into :610-653) a row of rates maps onto
// the board's columns.  A move that merged two tiles into a new masked one
// goes through dispatch (:682-708) and one of four lookups (update_mnt_osr_*,
// :747-1011); a move that made the target tile succeeds.  The reference
// memoises the column order per placement pattern (its match cache); the
// order is a function of that pattern, so it is computed here each time.
//
// Layout: one thread per board of a bucket (k_ad_solve), its row's rates and
// double accumulators in registers or per-lane scratch (row widths in tiers
// of 1, 8 and 64); buckets of layers i+1 / i+2 are found through a device
// table of 48 bucket descriptors (sorted boards, rates, rows, cols) and
// searched by binary search.  Every row index is checked against its
// bucket's width; an index outside it sets an error flag the host turns
// into an error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <dirent.h>
#include <memory>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

#include "egtb_ad.h"
#include "egtb_kernels.h"

namespace egtb {
namespace {

constexpr int kPermSlots = 256;  // (m << 4) + n
enum AdSolveErr : unsigned { kErrDerived = 1u, kErrColumn = 2u };

// One permutation table (m cells, n tiles): rows x cols bytes, rows whose
// first cell is f (first subset) and whose first two are (f, s) (pair
// subset), as build_permutation_subsets (BoardMaskerAD.cpp:360-407) makes.
struct PermSlotDev {
    uint32_t data, rows, cols;
    uint32_t first;      // 17 offsets in the word pool, then first_idx
    uint32_t first_idx;  // row indices
    uint32_t pair;       // 257 offsets
    uint32_t pair_idx;
};
struct AdPermDev {
    const PermSlotDev *slots;
    const uint8_t *bytes;
    const uint32_t *words;
};

template <typename T>
struct AdBucketDev {
    const uint64_t *ind;
    const T *rates;
    uint64_t rows;
    uint32_t cols, pad;
    const uint32_t *idx;  // range index (nullptr: plain binary search)
    uint64_t pre_mask, pre_val;
    uint32_t shift, bits;
};

template <typename T>
struct AdSolveCtx {
    AdSpecDev s;
    const uint8_t *comb;
    const uint16_t *mlut;  // masked row table
    const uint16_t *clut;  // classic row table
    AdPermDev perm;
    T zero, max_scale;
};

__device__ __forceinline__ uint32_t nib(uint64_t b, int k) { return (uint32_t)((b >> (4 * k)) & 0xF); }

// apply_sym_like (SymmetryUtils.h:9-29): 1 ReverseLR, 2 ReverseUD, 3 ReverseUL
// (transpose), 4 ReverseUR (anti-transpose), 5 Rotate180, 6 RotateL, 7 RotateR.
__device__ __forceinline__ uint64_t sym_like(uint64_t b, int idx) {
    switch (idx) {
        case 1: return mirror_lr(b);
        case 2: return mirror_ud(b);
        case 3: return transpose(b);
        case 4: return mirror_lr(mirror_ud(transpose(b)));
        case 5: return mirror_lr(mirror_ud(b));
        case 6: return mirror_ud(transpose(b));
        case 7: return mirror_lr(transpose(b));
        default: return b;
    }
}

// apply_sym_pair (Calculator.h:103-128): the canonical board and the first
// transform (in the order above) that gives it.
__device__ __forceinline__ uint64_t sym_pair(uint64_t b, uint32_t symm, int &idx) {
    idx = 0;
    if (symm == 0) return b;
    if (symm == 2) {
        const uint64_t t = transpose(b);
        if (b <= t) return b;
        idx = 3;
        return t;
    }
    if (symm == 3) {
        const uint64_t t = mirror_lr(b);
        if (b <= t) return b;
        idx = 1;
        return t;
    }
    uint64_t m = b;
    for (int k = 1; k < 8; ++k) {
        const uint64_t t = sym_like(b, k);
        if (t < m) {
            m = t;
            idx = k;
        }
    }
    return m;
}

// One move (0 left, 1 right, 2 up, 3 down: BoardMover::move_board(b, d + 1),
// CommonMover.h:126-134) through a row table.
__device__ __forceinline__ uint64_t move_dir(uint64_t b, int d, const uint16_t *lut) {
    switch (d) {
        case 0: return rows_left(b, lut);
        case 1: return mirror_lr(rows_left(mirror_lr(b), lut));
        case 2: return transpose(rows_left(transpose(b), lut));
        default: return transpose(mirror_lr(rows_left(mirror_lr(transpose(b)), lut)));
    }
}

// tile_sum_and_32k_count4 (BoardMaskerAD.cpp:219-257): pinned cells cleared;
// cells above 32 are large; pos_rank counts the 0xF cells above the highest
// non-0xF large cell (all of them when there is none).
struct Count4 {
    uint32_t total;
    int count;
    uint64_t pos_bitmap;
    int pos_rank;
    int merged_found;
    uint32_t merged_tile;
    bool success;
};

__device__ __forceinline__ Count4 count4(uint64_t b, const AdSpecDev &s) {
    b &= ~s.fixed_mask;
    Count4 r{0, 0, 0, 0, 0, 0, false};
    bool merged = false;
    for (int k = 15; k >= 0; --k) {
        const uint32_t v = nib(b, k);
        if (v > 5) {
            r.pos_bitmap |= 0xFull << (4 * k);
            ++r.count;
            if (v == 15) {
                if (!merged) ++r.pos_rank;
            } else {
                ++r.merged_found;
                r.merged_tile ^= v;
                merged = true;
            }
        } else if (v) {
            r.total += 1u << v;
        }
    }
    r.success = r.merged_tile == (uint32_t)s.target && r.merged_found == 1;
    return r;
}

// The large-tile combination for a large-tile sum and count (uint8 index
// coordinates, tiles_combination_view); nullptr when there is none.
__device__ __forceinline__ const uint8_t *comb_at(const uint8_t *comb, uint32_t lsum, int cnt) {
    const uint32_t idx = (uint32_t)(uint8_t)(lsum >> 6) + ((uint32_t)(uint8_t)cnt << 8);
    if (idx >= (uint32_t)kCombEntries) return nullptr;
    const uint8_t *c = comb + idx * kCombStride;
    return c[0] ? c : nullptr;
}

__device__ __forceinline__ uint32_t pinned_sum(const AdSpecDev &s) {
    return (uint32_t)(s.n_free + s.n_fixed) << 15;
}

// expand_ad's bucket key (BookSolverAD.cpp:1255-1286) with
// tile_sum_and_32k_count2's statistics; -128 drops the board.
__device__ __forceinline__ int bucket_key(uint64_t b, uint32_t bsum, const AdSpecDev &s, const uint8_t *comb) {
    b &= ~s.fixed_mask;
    uint32_t total = 0;
    int count = 0;
    for (int k = 0; k < 16; ++k) {
        const uint32_t v = nib(b, k);
        if (v > 5) ++count;
        else if (v) total += 1u << v;
    }
    if (count == s.n_free) return count;
    const uint8_t *c = comb_at(comb, bsum - total - pinned_sum(s), count - s.n_free);
    if (!c) return -128;
    if (count - s.n_free == 1) return count;
    if (c[0] > 1 && c[1] == c[2]) return (c[0] > 2 && c[1] == c[3]) ? (int)(int8_t)(count - 3 + 16) : -count;
    return count;
}

// unmask_board_into (BoardMaskerAD.cpp:445-500): the boards a masked board
// stands for, one per permutation-table row (tiles ascending over the 0xF
// cells, highest cell first); a combination with an equal pair keeps the
// rows whose second cell comes before the first.  Returns their number.
template <int DM>
__device__ int unmask(uint64_t bs, uint32_t ssum, const AdSpecDev &s, const uint8_t *comb, const AdPermDev &P,
                      uint64_t *D, unsigned &err) {
    const uint64_t b = bs & ~s.fixed_mask;
    uint8_t pos[16];
    int cnt = 0;
    uint32_t total = 0;
    for (int r = 3; r >= 0; --r) {
        uint32_t rs = 0;  // per-row sums are 16-bit in the reference's info table
        for (int i = 3; i >= 0; --i) {
            const int k = 4 * r + i;
            const uint32_t v = nib(b, k);
            if (v == 15) pos[cnt++] = (uint8_t)(4 * k);
            else if (v) rs += 1u << v;
        }
        total += rs & 0xFFFFu;
    }
    if (cnt == 0 || cnt == s.n_free) {
        D[0] = bs;
        return 1;
    }
    const uint8_t *c = comb_at(comb, ssum - total - pinned_sum(s), cnt - s.n_free);
    if (!c) return 0;
    const int m = cnt, n = cnt - s.n_free;
    if (m > 15 || n < 1 || n > 15) return 0;
    const PermSlotDev &sl = P.slots[(m << 4) + n];
    if (sl.rows == 0 || sl.cols == 0) return 0;
    const int nt = c[0];
    const bool eq = nt > 1 && c[1] == c[2];
    int out = 0;
    for (uint32_t r = 0; r < sl.rows; ++r) {
        const uint8_t *pr = P.bytes + sl.data + (size_t)r * sl.cols;
        if (eq && !(pr[1] < pr[0])) continue;
        uint64_t clear = 0, set = 0;
        for (int j = 0; j < nt && j < (int)sl.cols; ++j) {
            const int sh = pos[pr[j] < cnt ? pr[j] : 0];
            clear |= 0xFull << sh;
            set |= (uint64_t)c[1 + j] << sh;
        }
        if (out >= DM) {
            err |= kErrDerived;
            return out;
        }
        D[out++] = (bs & ~clear) | set;
    }
    return out;
}

// match_arr_into (BookSolverAD.cpp:610-653): rank of each board in a stable
// ascending sort.
__device__ __forceinline__ void rank_of(const uint64_t *D, int nd, uint32_t *R) {
    for (int i = 0; i < nd; ++i) {
        uint32_t r = 0;
        const uint64_t x = D[i];
        for (int j = 0; j < nd; ++j) r += (D[j] < x) || (D[j] == x && j < i);
        R[i] = r;
    }
}

// search_arr_ad (:466-478): the row of `board` in bucket `key`, or nullptr.
// A bucket's boards share the bits above its index window; the window's
// value selects the range to search (as the reference's AdaptiveIndex
// narrows the search of buckets over 100k boards, :413-430).
template <typename T>
__device__ const T *ad_find(const AdBucketDev<T> *L, int key, uint64_t board, uint32_t &cols) {
    if (key < kAdKeyMin || key > kAdKeyMax) return nullptr;
    const AdBucketDev<T> &B = L[key - kAdKeyMin];
    if (B.rows == 0 || B.cols == 0) return nullptr;
    uint64_t lo = 0, hi = B.rows;
    if (B.idx) {
        if ((board & B.pre_mask) != B.pre_val) return nullptr;
        const uint32_t v = (uint32_t)(board >> B.shift) & ((1u << B.bits) - 1u);
        lo = B.idx[v];
        hi = B.idx[v + 1];
    }
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (B.ind[mid] < board) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= B.rows || B.ind[lo] != board) return nullptr;
    cols = B.cols;
    return B.rates + lo * B.cols;
}

template <typename T>
__device__ __forceinline__ T tmax(T a, T b) {
    return a < b ? b : a;
}

// find_3x64_pos (:868-889)
__device__ __forceinline__ void find_3x64(uint64_t un, const AdSpecDev &s, int &pr64, int &pr128, int &prank,
                                          int &pos64) {
    pr64 = pr128 = prank = pos64 = 0;
    for (int sh = 60; sh >= 0; sh -= 4) {
        const uint32_t v = (uint32_t)((un >> sh) & 0xF);
        if (v == 15) {
            if (((s.fixed_mask >> sh) & 0xF) == 0) ++prank;
        } else if (v == 6) {
            pr64 = prank;
            pos64 = sh;
            ++prank;
        } else if (v == 7) {
            pr128 = prank;
            ++prank;
        }
    }
    prank = (int)(int8_t)prank;
}

// permutations_mapping_364 (:891-896)
__device__ __forceinline__ uint32_t perm364(int x, int y, int n) {
    if (x < y) return (uint32_t)(2 * n * x - x * x - 2 * x + y - 1);
    return (uint32_t)(2 * n * y - y * y - 3 * y + n + x - 2);
}

// ---- one rate per board (derive_size 1) -----------------------------------

// update_mnt_osr_ad (:956-1011)
template <typename T>
__device__ T upd_mnt(const AdBucketDev<T> *L, uint64_t board, uint64_t un, T osr, uint32_t ssum,
                     const AdSolveCtx<T> &c, unsigned &err) {
    const Count4 st = count4(un, c.s);
    if (st.success) return c.max_scale;
    const uint8_t *cb = comb_at(c.comb, ssum - st.total - pinned_sum(c.s), st.count - c.s.n_free);
    if (!cb) return osr;
    if (st.merged_found == 2) return osr;
    int bk = st.count;
    if (cb[0] > 1 && cb[1] == cb[2]) {
        if (cb[0] > 2 && cb[1] == cb[3]) {
            bk = (int)(int8_t)(st.count + 16 - 3);
            board = un;
        } else {
            bk = (int)(int8_t)(-st.count);
            const uint64_t all = (uint64_t)cb[1] * 0x1111111111111111ull;
            board = (board & ~st.pos_bitmap) | (all & st.pos_bitmap);
        }
    }
    uint32_t cols = 0;
    const T *row = ad_find(L, bk, board, cols);
    if (!row) return osr;
    const uint32_t col = (bk > 0 && bk <= 15) ? (uint32_t)st.pos_rank : 0u;
    if (col >= cols) {
        err |= kErrColumn;
        return osr;
    }
    return tmax(osr, row[col]);
}

// update_mnt_osr_364_ad (:898-916)
template <typename T>
__device__ T upd_364(const AdBucketDev<T> *L, uint64_t board, uint64_t un, T osr, const AdSolveCtx<T> &c,
                     unsigned &err) {
    int pr64, pr128, prank, pos64;
    find_3x64(un, c.s, pr64, pr128, prank, pos64);
    board |= 0xFull << pos64;
    uint32_t cols = 0;
    const T *row = ad_find(L, prank, board, cols);
    if (!row) return osr;
    const uint32_t col = perm364(pr64, pr128, prank);
    if (col >= cols) {
        err |= kErrColumn;
        return osr;
    }
    return tmax(osr, row[col]);
}

// solve_optimal_success_rate (:1178-1241)
template <typename T>
__device__ T sosr_one(uint64_t b, uint64_t v, int pos, uint32_t ssum, int key, const AdBucketDev<T> *L,
                      const AdSolveCtx<T> &c, unsigned &err) {
    const uint64_t bs = b | (v << (4 * pos));
    uint64_t mv[4];
    move4(bs, c.mlut, mv);
    const int f0 = f_count(bs);
    const bool fh = f_count(mv[0]) > f0, fv = f_count(mv[2]) > f0;  // left's flag for left/right, up's for up/down
    T osr = c.zero;
    for (int d = 0; d < 4; ++d) {
        const uint64_t m = mv[d];
        if (m == bs || !ad_is_pattern(m, c.s)) continue;
        int si;
        const uint64_t canon = sym_pair(m, c.s.symm, si);
        if (d < 2 ? fh : fv) {
            const uint64_t un = sym_like(move_dir(bs, d, c.clut), si);
            osr = key < 16 ? upd_mnt(L, canon, un, osr, ssum, c, err) : upd_364(L, canon, un, osr, c, err);
        } else {
            uint32_t cols = 0;
            const T *row = ad_find(L, key, canon, cols);
            if (row) osr = tmax(osr, row[0]);
        }
    }
    return osr;
}

// ---- a row of rates per board ---------------------------------------------

template <typename T>
__device__ __forceinline__ void raise_row(const T *row, uint32_t cols, const uint32_t *idx, int nd, T *osr,
                                          unsigned &err) {
    for (int i = 0; i < nd; ++i) {
        if (idx[i] >= cols) {
            err |= kErrColumn;
            return;
        }
        const T x = row[idx[i]];
        if (x > osr[i]) osr[i] = x;
    }
}

// row[subset[R[i]]] into osr[i] (update_mnt_osr_ad_arr2 / _364_arr_ad)
template <typename T>
__device__ __forceinline__ void raise_subset(const T *row, uint32_t cols, const uint32_t *subset, uint32_t size,
                                             const uint32_t *R, int nd, T *osr, unsigned &err) {
    if (size < (uint32_t)nd) return;
    for (int i = 0; i < nd; ++i) {
        if (R[i] >= size || subset[R[i]] >= cols) {
            err |= kErrColumn;
            return;
        }
        const T x = row[subset[R[i]]];
        if (x > osr[i]) osr[i] = x;
    }
}

// solve_optimal_success_rate_arr_into (:1013-1176) for one spawn: osr[0..d)
template <typename T, int DM>
__device__ void sosr_row(uint64_t b, uint64_t v, int pos, uint32_t ssum, int key, int d, const AdBucketDev<T> *L,
                         const AdSolveCtx<T> &c, T *osr, uint64_t *D, uint32_t *R, uint64_t *mb, uint32_t *R2,
                         uint8_t *mi, unsigned &err) {
    const AdSpecDev &s = c.s;
    const uint64_t bs = b | (v << (4 * pos));
    for (int j = 0; j < d; ++j) osr[j] = c.zero;
    uint64_t mv[4];
    move4(bs, c.mlut, mv);
    const int f0 = f_count(bs);
    const bool fh = f_count(mv[0]) > f0, fv = f_count(mv[2]) > f0;
    for (int dir = 0; dir < 4; ++dir) {
        const uint64_t m = mv[dir];
        if (m == bs || !ad_is_pattern(m, s)) continue;
        int si;
        const uint64_t canon = sym_pair(m, s.symm, si);
        // the placements of bs moved and in canonical orientation, and their
        // column order (process_derived_into + match_arr_into), made where a
        // lookup needs them
        int nd = -1;
        auto placements = [&]() {
            if (nd >= 0) return;
            nd = unmask<DM>(bs, ssum, s, c.comb, c.perm, D, err);
            if (nd > d) {
                err |= kErrDerived;
                nd = d;
            }
            for (int j = 0; j < nd; ++j) D[j] = sym_like(move_dir(D[j], dir, c.clut), si);
            rank_of(D, nd, R);
        };
        if (dir < 2 ? fh : fv) {
            const uint64_t un = sym_like(move_dir(bs, dir, c.clut), si);
            if (key > 15) {
                // update_mnt_osr_364_arr_ad (:918-954)
                placements();
                int pr64, pr128, prank, pos64;
                find_3x64(un, s, pr64, pr128, prank, pos64);
                uint32_t cols = 0;
                const T *row = ad_find(L, prank, canon | (0xFull << pos64), cols);
                const int pm = prank, pn = prank - s.n_free;
                if (row && pm <= 15 && pn >= 1 && pn <= 15) {
                    const PermSlotDev &sl = c.perm.slots[(pm << 4) + pn];
                    if (sl.rows && sl.cols > 1) {
                        const uint32_t code = ((uint32_t)pr64 << 4) + (uint32_t)pr128;
                        const uint32_t *off = c.perm.words + sl.pair;
                        raise_subset(row, cols, c.perm.words + sl.pair_idx + off[code], off[code + 1] - off[code], R,
                                     nd, osr, err);
                    }
                }
                continue;
            }
            // dispatch_mnt_osr_ad_arr (:682-708)
            const Count4 st = count4(un, s);
            if (st.success) {
                for (int j = 0; j < d; ++j) osr[j] = c.max_scale;
                continue;
            }
            const uint8_t *cb = comb_at(c.comb, ssum - st.total - pinned_sum(s), st.count - s.n_free);
            if (!cb) continue;
            const int nt = cb[0];
            if (nt > 2 && cb[1] == cb[2] && cb[1] == cb[3]) {
                // update_mnt_osr_ad_arr3 (:747-769)
                placements();
                uint32_t cols = 0;
                const T *row = ad_find(L, (int)(int8_t)(st.count - 3 + 16), un, cols);
                if (row) raise_row(row, cols, R, nd, osr, err);
            } else if (nt > 1 && cb[1] == cb[2]) {
                // update_mnt_osr_ad_arr1 (:807-866): each other large cell
                // takes the pair's tile; the placements with the merged tile
                // on it map onto that board's row
                placements();
                uint8_t lp[16];
                int nl = 0;
                for (uint64_t pm = st.pos_bitmap; pm && nl < 16;) {
                    const int msb = 63 - __builtin_clzll(pm);
                    const int sh = msb & ~3;
                    lp[nl++] = (uint8_t)sh;
                    pm &= ~(0xFull << sh);
                }
                for (int j = 0; j < st.count && j < nl; ++j) {
                    if (j == st.pos_rank) continue;
                    const int p = lp[j];
                    const uint64_t ubj = (un & ~(0xFull << p)) | ((uint64_t)cb[1] << p);
                    int si2;
                    const uint64_t cbj = sym_pair(ubj, s.symm, si2);
                    uint32_t cols = 0;
                    const T *row = ad_find(L, -st.count, cbj, cols);
                    if (!row) continue;
                    int nm = 0;
                    for (int q = 0; q < nd; ++q)
                        if (((D[q] >> p) & 0xF) == st.merged_tile) {
                            mi[nm] = (uint8_t)q;
                            mb[nm] = sym_like(D[q], si2);
                            ++nm;
                        }
                    if (!nm) continue;
                    rank_of(mb, nm, R2);
                    for (int q = 0; q < nm; ++q) {
                        if (R2[q] >= cols) {
                            err |= kErrColumn;
                            break;
                        }
                        const T x = row[R2[q]];
                        if (x > osr[mi[q]]) osr[mi[q]] = x;
                    }
                }
            } else {
                // update_mnt_osr_ad_arr2 (:771-805)
                placements();
                uint32_t cols = 0;
                const T *row = ad_find(L, st.count, canon, cols);
                const int pm = st.count, pn = st.count - s.n_free;
                if (row && pm <= 15 && pn >= 1 && pn <= 15 && st.pos_rank < 16) {
                    const PermSlotDev &sl = c.perm.slots[(pm << 4) + pn];
                    if (sl.rows) {
                        const uint32_t *off = c.perm.words + sl.first;
                        raise_subset(row, cols, c.perm.words + sl.first_idx + off[st.pos_rank],
                                     off[st.pos_rank + 1] - off[st.pos_rank], R, nd, osr, err);
                    }
                }
            }
            continue;
        }
        // update_osr_ad_arr (:710-730): the same bucket of the next layer
        placements();
        uint32_t cols = 0;
        const T *row = ad_find(L, key, canon, cols);
        if (row) raise_row(row, cols, R, nd, osr, err);
    }
}

// recalculate_ad (:1321-1505) over one bucket of layer i: rows [0, rows),
// cols rates each, into out.  The arithmetic is the reference's as its
// -Ofast build (fp contraction, reassociation, reciprocal math) compiles it:
// one rate per board: per empty cell sum += fma(rate2, 1 - p4, rate4 * p4),
// then sum / empty; a row of rates: per empty cell and column
// sum = fma(rate2, 1 - p4, sum), then sum = fma(rate4, p4, sum), and at the
// end sum * (1 / empty).
template <typename T, int DM>
__global__ __launch_bounds__(256) void k_ad_solve(const uint64_t *__restrict__ ind, uint64_t rows, uint32_t cols,
                                                  int key, T *__restrict__ out, const AdBucketDev<T> *L1,
                                                  const AdBucketDev<T> *L2, uint32_t bsum, AdSolveCtx<T> c,
                                                  double w2, double w4, unsigned *errp,
                                                  unsigned long long *maxbits) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long mx = 0;  // bits of the row's largest rate (rates are >= 0)
    auto note = [&](T v) {
        unsigned long long x = 0;
        __builtin_memcpy(&x, &v, sizeof(T));
        mx = x > mx ? x : mx;
    };
    if (r < rows) {
    const uint64_t b = ind[r];
    unsigned err = 0;
    const uint32_t em = empty_cells(b);
    const int empty = __popc(em);
    if (DM == 1) {
        double sp = 0.0;
        for (uint32_t m = em; m;) {
            const int pos = __builtin_ctz(m);
            m &= m - 1;
            const T s2 = sosr_one(b, 1, pos, bsum + 2u, key, L1, c, err);
            const T s4 = sosr_one(b, 2, pos, bsum + 4u, key, L2, c, err);
            sp = __dadd_rn(sp, __fma_rn((double)s2, w2, __dmul_rn((double)s4, w4)));
        }
        const T v = empty == 0 ? c.zero : (T)__ddiv_rn(sp, (double)empty);
        out[r] = v;
        note(v);
    } else {
        const int d = (int)cols;
        double sp[DM];
        T osr[DM];
        uint64_t D[DM], mb[DM];
        uint32_t R[DM], R2[DM];
        uint8_t mi[DM];
        for (int j = 0; j < d; ++j) sp[j] = 0.0;
        for (uint32_t m = em; m;) {
            const int pos = __builtin_ctz(m);
            m &= m - 1;
            sosr_row<T, DM>(b, 1, pos, bsum + 2u, key, d, L1, c, osr, D, R, mb, R2, mi, err);
            for (int j = 0; j < d; ++j) sp[j] = __fma_rn((double)osr[j], w2, sp[j]);
            sosr_row<T, DM>(b, 2, pos, bsum + 4u, key, d, L2, c, osr, D, R, mb, R2, mi, err);
            for (int j = 0; j < d; ++j) sp[j] = __fma_rn((double)osr[j], w4, sp[j]);
        }
        T *o = out + r * cols;
        const double inv = empty == 0 ? 0.0 : __ddiv_rn(1.0, (double)empty);
        for (int j = 0; j < d; ++j) {
            const T v = empty == 0 ? c.zero : (T)__dmul_rn(sp[j], inv);
            o[j] = v;
            note(v);
        }
    }
    if (err) atomicOr(errp, err);
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(mx, off, 64);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0 && mx) atomicMax(maxbits, mx);
}

// the range index of a sorted bucket: idx[v] = first board whose window
// value is >= v (boards all share pre_val above the window), idx[2^bits] = n
__global__ void k_ad_range_index(const uint64_t *__restrict__ keys, uint64_t n, uint64_t pre_val, uint32_t shift,
                                 uint32_t bits, uint32_t *__restrict__ idx) {
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nb = 1ull << bits;
    if (v > nb) return;
    if (v == nb) {
        idx[v] = (uint32_t)n;
        return;
    }
    const uint64_t target = pre_val | (v << shift);
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (keys[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    idx[v] = (uint32_t)lo;
}

// bucket key of every board of a raw layer (int8; -128 = dropped) and the
// bucket sizes
__global__ void k_ad_bucket_keys(const uint64_t *__restrict__ keys, uint64_t n, uint32_t bsum, AdSpecDev s,
                                 const uint8_t *__restrict__ comb, int8_t *__restrict__ out,
                                 unsigned long long *__restrict__ hist) {
    __shared__ unsigned h[kAdKeys];
    for (int i = threadIdx.x; i < kAdKeys; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        int k = bucket_key(keys[i], bsum, s, comb);
        if (k < kAdKeyMin || k > kAdKeyMax) k = -128;
        out[i] = (int8_t)k;
        if (k != -128) atomicAdd(&h[k - kAdKeyMin], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kAdKeys; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], (unsigned long long)h[i]);
}

__global__ void k_ad_key_flags(const int8_t *__restrict__ k, uint64_t n, int key, uint32_t *__restrict__ flags) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = k[i] == (int8_t)key ? 1u : 0u;
}

// rows of a bucket with some rate above thr (remove_died_ad, :495-536)
template <typename T>
__global__ void k_ad_live_rows(const T *__restrict__ rates, uint64_t rows, uint32_t cols, T thr,
                               uint32_t *__restrict__ flags) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    uint32_t keep = 0;
    for (uint32_t j = 0; j < cols; ++j) keep |= rates[r * cols + j] > thr ? 1u : 0u;
    flags[r] = keep;
}

// stable compaction of the flagged rows (boards, and cols rates each when
// rates is not NULL) to out positions offs[r]; count gets the kept rows
template <typename T>
__global__ void k_ad_scatter_rows(const uint64_t *__restrict__ ind, const T *__restrict__ rates, uint64_t rows,
                                  uint32_t cols, const uint32_t *__restrict__ flags,
                                  const uint32_t *__restrict__ offs, uint64_t *__restrict__ out_ind,
                                  T *__restrict__ out_rates, unsigned long long *__restrict__ count) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    if (r == rows - 1) *count = (unsigned long long)offs[r] + flags[r];
    if (!flags[r]) return;
    const uint64_t o = offs[r];
    out_ind[o] = ind[r];
    if (rates)
        for (uint32_t j = 0; j < cols; ++j) out_rates[o * cols + j] = rates[r * cols + j];
}

template <typename T>
struct RateKind;
template <>
struct RateKind<float> {
    static float max_scale() { return 1.0f; }
};
template <>
struct RateKind<double> {
    static double max_scale() { return 1.0; }
};
template <>
struct RateKind<uint32_t> {
    static uint32_t max_scale() { return 4000000000u; }
};
template <>
struct RateKind<uint64_t> {
    static uint64_t max_scale() { return 1600000000000000000ull; }
};

const uint32_t kFact[19] = {1u, 1u, 2u, 6u, 24u, 120u, 720u, 5040u, 40320u, 362880u,
                            3628800u, 39916800u, 1u, 1u, 1u, 1u, 1u, 1u, 1u};

// derive_size_for_bucket (:291-299)
uint32_t derive_size(int key, int n_free) {
    const int a = key < 0 ? -key - 2 : key > 15 ? key - 16 : key;
    if (a < 0 || a > 18 || n_free < 0 || n_free > 18) return 0;
    return kFact[a] / kFact[n_free];
}

// Permutation tables for n = 1 .. target-5 tiles over m = n + n_free cells
// (init_masker, BoardMaskerAD.cpp:514-521): every ordered choice of n of the
// m cells, rows sorted by the board that puts tile j on cell row[j] (cell
// i at nibble m-1-i, other cells 0xF), with the first / pair subsets.
struct PermHost {
    std::vector<PermSlotDev> slots = std::vector<PermSlotDev>(kPermSlots, PermSlotDev{0, 0, 0, 0, 0, 0, 0});
    std::vector<uint8_t> bytes;
    std::vector<uint32_t> words;
};

PermHost make_perms(int target, int n_free) {
    PermHost P;
    P.bytes.push_back(0);
    for (int n = 1; n < target - 4; ++n) {
        const int m = n + n_free;
        if (m > 15 || n > 15) break;
        uint64_t count = 1;
        for (int k = 0; k < n; ++k) count *= (uint64_t)(m - k);
        if (count > (1u << 16)) break;  // rows wider than the solver's rows are never needed
        std::vector<std::pair<uint64_t, std::vector<uint8_t>>> rows;
        std::vector<uint8_t> cur;
        std::vector<bool> used(m, false);
        auto dfs = [&](auto &&self) -> void {
            if ((int)cur.size() == n) {
                uint64_t clear = 0, set = 0;
                for (int j = 0; j < n; ++j) {
                    const int sh = (m - 1 - cur[j]) * 4;
                    clear |= 0xFull << sh;
                    set |= (uint64_t)j << sh;
                }
                rows.push_back({~clear | set, cur});
                return;
            }
            for (int i = 0; i < m; ++i) {
                if (used[i]) continue;
                used[i] = true;
                cur.push_back((uint8_t)i);
                self(self);
                cur.pop_back();
                used[i] = false;
            }
        };
        dfs(dfs);
        std::sort(rows.begin(), rows.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        PermSlotDev sl{};
        sl.data = (uint32_t)P.bytes.size();
        sl.rows = (uint32_t)rows.size();
        sl.cols = (uint32_t)n;
        for (auto &r : rows) P.bytes.insert(P.bytes.end(), r.second.begin(), r.second.end());
        // first subsets
        std::vector<uint32_t> fo(17, 0), po(257, 0);
        for (auto &r : rows) {
            ++fo[r.second[0] + 1];
            if (n > 1) ++po[((size_t)r.second[0] << 4) + r.second[1] + 1];
        }
        for (int i = 0; i < 16; ++i) fo[i + 1] += fo[i];
        for (int i = 0; i < 256; ++i) po[i + 1] += po[i];
        std::vector<uint32_t> fi(rows.size()), pi(n > 1 ? rows.size() : 0);
        std::vector<uint32_t> fw(fo.begin(), fo.end() - 1), pw(po.begin(), po.end() - 1);
        for (size_t r = 0; r < rows.size(); ++r) {
            fi[fw[rows[r].second[0]]++] = (uint32_t)r;
            if (n > 1) pi[pw[((size_t)rows[r].second[0] << 4) + rows[r].second[1]]++] = (uint32_t)r;
        }
        sl.first = (uint32_t)P.words.size();
        P.words.insert(P.words.end(), fo.begin(), fo.end());
        sl.first_idx = (uint32_t)P.words.size();
        P.words.insert(P.words.end(), fi.begin(), fi.end());
        sl.pair = (uint32_t)P.words.size();
        P.words.insert(P.words.end(), po.begin(), po.end());
        sl.pair_idx = (uint32_t)P.words.size();
        P.words.insert(P.words.end(), pi.begin(), pi.end());
        P.slots[(m << 4) + n] = sl;
    }
    if (P.words.empty()) P.words.push_back(0);
    return P;
}

bool make_dir(const std::string &p) { return ::mkdir(p.c_str(), 0777) == 0 || errno == EEXIST; }

bool is_dir(const std::string &p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

std::vector<uint8_t> read_bytes(const std::string &p) {
    FILE *f = std::fopen(p.c_str(), "rb");
    if (!f) return {};
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<uint8_t> v(sz > 0 ? (size_t)sz : 0);
    const size_t got = v.empty() ? 0 : std::fread(v.data(), 1, v.size(), f);
    std::fclose(f);
    if (got != v.size()) throw Error(EGTB_ERR_IO, "short read of " + p);
    return v;
}

std::string now_string() {
    const std::time_t t = std::time(nullptr);
    std::tm lt{};
    localtime_r(&t, &lt);
    char buf[32];
    std::strftime(buf, sizeof buf, "%Y-%m-%d %H:%M:%S", &lt);
    return buf;
}

std::string fmt_g(double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%g", v);
    return buf;
}

template <typename T>
struct AdSolver {
    hipStream_t s;
    int cus;
    AdSolveCtx<T> ctx;
    DBuf err, hist, acc, flags, offs, part, keys8, perm_slots, perm_bytes, perm_words;
    HostStage stage;

    uint64_t read_u64(const void *p) {
        uint64_t v = 0;
        HIPCHK(hipMemcpyAsync(&v, p, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return v;
    }

    // the device table of a finished layer's 48 bucket descriptors, kept
    // with the layer (its host copy too, so the copy needs no wait)
    void describe(AdBookLayer &L) {
        L.desc_host.resize(sizeof(AdBucketDev<T>) * kAdKeys);
        auto *h = (AdBucketDev<T> *)L.desc_host.data();
        for (int k = 0; k < kAdKeys; ++k) {
            const AdBucket &B = L.b[k];
            h[k] = AdBucketDev<T>{(const uint64_t *)B.ind.p, (const T *)B.rates.p, B.rows, B.cols, 0,
                                  (const uint32_t *)B.idx.p, B.pre_mask, B.pre_val, B.shift, B.bits};
        }
        L.desc.ensure(L.desc_host.size(), s);
        HIPCHK(hipMemcpyAsync(L.desc.p, L.desc_host.data(), L.desc_host.size(), hipMemcpyHostToDevice, s));
    }

    // remove_died_ad: rows with no rate above thr leave the layer.  Every
    // bucket is compacted into new buffers (the flag and offset buffers are
    // reused in stream order), its kept count into acc[k]; one read at the end.
    void prune(AdBookLayer &L, T thr) {
        DBuf ni[kAdKeys], nr[kAdKeys];
        bool any = false;
        acc.ensure(kAdKeys * 8, s);
        for (int k = 0; k < kAdKeys; ++k) {
            AdBucket &B = L.b[k];
            if (B.rows == 0 || B.cols == 0) continue;
            const uint64_t rows = B.rows;
            flags.ensure(rows * 4 + 4, s);
            offs.ensure(rows * 4 + 4, s);
            part.ensure(scan_partials_needed(rows) * 4 + 64, s);
            hipLaunchKernelGGL(k_ad_live_rows<T>, dim3(grid_for(rows, 256, 1u << 30)), dim3(256), 0, s,
                               (const T *)B.rates.p, rows, B.cols, thr, (uint32_t *)flags.p);
            HIPCHK(hipGetLastError());
            launch_scan_u32(s, (const uint32_t *)flags.p, rows, (uint32_t *)offs.p, (uint32_t *)part.p);
            ni[k].alloc(rows * 8 + 8, s);
            nr[k].alloc(rows * B.cols * sizeof(T) + 8, s);
            hipLaunchKernelGGL(k_ad_scatter_rows<T>, dim3(grid_for(rows, 256, 1u << 30)), dim3(256), 0, s,
                               (const uint64_t *)B.ind.p, (const T *)B.rates.p, rows, B.cols,
                               (const uint32_t *)flags.p, (const uint32_t *)offs.p, (uint64_t *)ni[k].p,
                               (T *)nr[k].p, (unsigned long long *)acc.p + k);
            HIPCHK(hipGetLastError());
            any = true;
        }
        if (!any) return;
        uint64_t kept[kAdKeys];
        HIPCHK(hipMemcpyAsync(kept, acc.p, sizeof kept, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (int k = 0; k < kAdKeys; ++k) {
            AdBucket &B = L.b[k];
            if (!ni[k].p) continue;
            B.rows = kept[k];
            B.ind = std::move(ni[k]);
            B.rates = std::move(nr[k]);
            B.idx.release();
            if (B.rows == 0) B.cols = 0;
        }
    }

    // expand_ad: the raw layer's boards into buckets (sorted within each)
    void bucketize(const uint64_t *keys, uint64_t n, uint32_t bsum, int n_free, AdBookLayer &L) {
        if (n == 0) return;
        keys8.ensure(n + 8, s);
        hist.ensure(kAdKeys * 8, s);
        HIPCHK(hipMemsetAsync(hist.p, 0, kAdKeys * 8, s));
        hipLaunchKernelGGL(k_ad_bucket_keys, dim3(grid_for(n, 256, (uint32_t)cus * 16)), dim3(256), 0, s, keys, n,
                           bsum, ctx.s, ctx.comb, (int8_t *)keys8.p, (unsigned long long *)hist.p);
        HIPCHK(hipGetLastError());
        uint64_t h[kAdKeys];
        HIPCHK(hipMemcpyAsync(h, hist.p, sizeof h, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (int k = 0; k < kAdKeys; ++k) {
            if (!h[k]) continue;
            AdBucket &B = L.b[k];
            // flag the bucket's boards, then compact the layer into it
            flags.ensure(n * 4 + 4, s);
            hipLaunchKernelGGL(k_ad_key_flags, dim3(grid_for(n, 256, 1u << 30)), dim3(256), 0, s,
                               (const int8_t *)keys8.p, n, k + kAdKeyMin, (uint32_t *)flags.p);
            HIPCHK(hipGetLastError());
            offs.ensure(n * 4 + 4, s);
            part.ensure(scan_partials_needed(n) * 4 + 64, s);
            launch_scan_u32(s, (const uint32_t *)flags.p, n, (uint32_t *)offs.p, (uint32_t *)part.p);
            B.ind.alloc(h[k] * 8 + 8, s);
            acc.ensure(64, s);
            hipLaunchKernelGGL(k_ad_scatter_rows<T>, dim3(grid_for(n, 256, 1u << 30)), dim3(256), 0, s, keys,
                               (const T *)nullptr, n, 0u, (const uint32_t *)flags.p, (const uint32_t *)offs.p,
                               (uint64_t *)B.ind.p, (T *)nullptr, (unsigned long long *)acc.p);
            HIPCHK(hipGetLastError());
            B.rows = h[k];
            B.cols = derive_size(k + kAdKeyMin, n_free);
            if (B.cols) B.rates.alloc(B.rows * B.cols * sizeof(T) + 8, s);
        }
    }

    void solve_bucket(int k, AdBucket &B, const DBuf &L1, const DBuf &L2, uint32_t bsum, double w2, double w4) {
        if (B.rows == 0 || B.cols == 0) return;
        const dim3 g(grid_for(B.rows, 256, 1u << 30)), t(256);
        const int key = k + kAdKeyMin;
        const auto *l1 = (const AdBucketDev<T> *)L1.p;
        const auto *l2 = (const AdBucketDev<T> *)L2.p;
        const auto *ind = (const uint64_t *)B.ind.p;
        T *out = (T *)B.rates.p;
        unsigned *e = (unsigned *)err.p;
        unsigned long long *mx = (unsigned long long *)err.p + 1;
        if (B.cols == 1)
            hipLaunchKernelGGL((k_ad_solve<T, 1>), g, t, 0, s, ind, B.rows, B.cols, key, out, l1, l2, bsum, ctx, w2,
                               w4, e, mx);
        else if (B.cols <= 8)
            hipLaunchKernelGGL((k_ad_solve<T, 8>), g, t, 0, s, ind, B.rows, B.cols, key, out, l1, l2, bsum, ctx, w2,
                               w4, e, mx);
        else if (B.cols <= 64)
            hipLaunchKernelGGL((k_ad_solve<T, 64>), g, t, 0, s, ind, B.rows, B.cols, key, out, l1, l2, bsum, ctx,
                               w2, w4, e, mx);
        else
            throw Error(EGTB_ERR_ARG, "advanced solve: rows of " + std::to_string(B.cols) +
                                          " rates (bucket " + std::to_string(key) + ") exceed the 64 supported");
        HIPCHK(hipGetLastError());
    }

    // the range index of every bucket of a finished layer with more than
    // kIndexMin boards: a window of about log2(rows) - 1 bits right below the
    // prefix the sorted boards share (that of the first and the last)
    void index_layer(AdBookLayer &L) {
        constexpr uint64_t kIndexMin = 256;
        uint64_t ends[kAdKeys][2] = {};
        bool any = false;
        for (int k = 0; k < kAdKeys; ++k) {
            const AdBucket &B = L.b[k];
            if (B.rows < kIndexMin || B.cols == 0 || B.rows > 0xFFFFFFFFull) continue;
            HIPCHK(hipMemcpyAsync(&ends[k][0], B.ind.p, 8, hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(&ends[k][1], (const uint64_t *)B.ind.p + (B.rows - 1), 8, hipMemcpyDeviceToHost, s));
            any = true;
        }
        if (!any) return;
        HIPCHK(hipStreamSynchronize(s));
        for (int k = 0; k < kAdKeys; ++k) {
            AdBucket &B = L.b[k];
            B.idx.release();
            if (B.rows < kIndexMin || B.cols == 0 || B.rows > 0xFFFFFFFFull) continue;
            const uint64_t diff = ends[k][0] ^ ends[k][1];
            if (!diff) continue;
            const int top = 63 - __builtin_clzll(diff);
            int want = 0;
            while ((1ull << (want + 1)) <= B.rows) ++want;  // floor(log2(rows))
            const int bits = std::max(1, std::min({want - 1, 22, top + 1}));
            B.bits = (uint32_t)bits;
            B.shift = (uint32_t)(top + 1 - bits);
            B.pre_mask = top == 63 ? 0ull : ~((2ull << top) - 1ull);
            B.pre_val = ends[k][0] & B.pre_mask;
            B.idx.alloc(((1ull << bits) + 1) * 4 + 8, s);
            hipLaunchKernelGGL(k_ad_range_index, dim3(grid_for((1ull << bits) + 1, 256, 1u << 30)), dim3(256), 0, s,
                               (const uint64_t *)B.ind.p, B.rows, B.pre_val, B.shift, B.bits, (uint32_t *)B.idx.p);
            HIPCHK(hipGetLastError());
        }
    }

    // dict_tofile (:320-352)
    void write_layer(const std::string &folder, const AdBookLayer &L) {
        if (!make_dir(folder)) throw Error(EGTB_ERR_IO, "cannot create " + folder);
        for (int k = 0; k < kAdKeys; ++k) {
            const AdBucket &B = L.b[k];
            const std::string base = folder + "/" + std::to_string(k + kAdKeyMin);
            if (B.rows == 0) {
                ::unlink((base + ".i").c_str());
                ::unlink((base + ".b").c_str());
                continue;
            }
            const uint64_t bi = B.rows * 8, br = B.rows * B.cols * sizeof(T);
            stage.ensure(std::max(bi, br) + 8);
            HIPCHK(hipMemcpyAsync(stage.p, B.ind.p, bi, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            write_file_atomic(base + ".i", stage.p, bi);
            if (br) HIPCHK(hipMemcpyAsync(stage.p, B.rates.p, br, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            write_file_atomic(base + ".b", stage.p, br);
        }
    }

    // dict_fromfile (:354-411)
    void read_layer(const std::string &folder, AdBookLayer &L) {
        if (!is_dir(folder)) throw Error(EGTB_ERR_IO, "Missing AD solve bucket folder: " + folder);
        DIR *d = ::opendir(folder.c_str());
        if (!d) throw Error(EGTB_ERR_IO, "cannot open " + folder);
        std::vector<int> keys;
        while (dirent *e = ::readdir(d)) {
            const std::string nm = e->d_name;
            if (nm.size() > 3 && nm.compare(nm.size() - 3, 3, ".zi") == 0) {
                ::closedir(d);
                throw Error(EGTB_ERR_IO, "compressed bucket index " + folder + "/" + nm + " is not supported");
            }
            if (nm.size() < 3 || nm.compare(nm.size() - 2, 2, ".i") != 0) continue;
            try {
                keys.push_back(std::stoi(nm.substr(0, nm.size() - 2)));
            } catch (const std::exception &) {
                ::closedir(d);
                throw Error(EGTB_ERR_IO, "unexpected file " + folder + "/" + nm);
            }
        }
        ::closedir(d);
        for (int key : keys) {
            if (key < kAdKeyMin || key > kAdKeyMax) throw Error(EGTB_ERR_IO, "bucket key out of range in " + folder);
            AdBucket &B = L.b[key - kAdKeyMin];
            const std::string base = folder + "/" + std::to_string(key);
            const std::vector<uint64_t> ind = read_raw(base + ".i");
            const std::vector<uint8_t> rb = read_bytes(base + ".b");
            B.rows = ind.size();
            if (B.rows == 0) continue;
            B.cols = rb.empty() ? 0 : (uint32_t)(rb.size() / sizeof(T) / B.rows);
            B.ind.alloc(B.rows * 8 + 8, s);
            HIPCHK(hipMemcpyAsync(B.ind.p, ind.data(), B.rows * 8, hipMemcpyHostToDevice, s));
            if (B.cols) {
                B.rates.alloc(B.rows * B.cols * sizeof(T) + 8, s);
                HIPCHK(hipMemcpyAsync(B.rates.p, rb.data(), B.rows * B.cols * sizeof(T), hipMemcpyHostToDevice, s));
            }
            HIPCHK(hipStreamSynchronize(s));  // the host vectors leave scope
        }
    }
};

template <typename T>
void ad_solve_t(const uint64_t *seeds, size_t n_seeds, const egtb_ad_pattern &pat, const egtb_ad_options &o,
                const egtb_ad_solve_options &so, AdResult &res) {
    if (!(so.spawn_rate4 >= 0.0 && so.spawn_rate4 <= 1.0)) throw Error(EGTB_ERR_ARG, "spawn_rate4 must be in [0, 1]");
    if (!(so.deletion_threshold >= 0.0 && so.deletion_threshold < 1.0))
        throw Error(EGTB_ERR_ARG, "deletion_threshold must be in [0, 1)");
    const AdSpecDev sp = make_ad_spec(pat);
    if (o.steps < 3) throw Error(EGTB_ERR_ARG, "an advanced build needs at least 3 steps");
    egtb_ad_options fo = o;
    fo.keep_layers = 1;
    ad_generate(seeds, n_seeds, pat, fo, res);

    const int dev = o.device;
    DeviceCtx &C = device_ctx(dev);
    HIPCHK(hipSetDevice(dev));
    const AdDevice &D = ad_device(dev);
    AdSolver<T> S;
    S.s = C.stream;
    S.cus = C.cus;
    const hipStream_t s = S.s;
    const auto t0 = std::chrono::steady_clock::now();
    const PermHost P = make_perms(pat.target, pat.n_free_32k);
    S.perm_slots.alloc(P.slots.size() * sizeof(PermSlotDev), s);
    S.perm_bytes.alloc(P.bytes.size() + 8, s);
    S.perm_words.alloc(P.words.size() * 4 + 8, s);
    HIPCHK(hipMemcpyAsync(S.perm_slots.p, P.slots.data(), P.slots.size() * sizeof(PermSlotDev),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(S.perm_bytes.p, P.bytes.data(), P.bytes.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(S.perm_words.p, P.words.data(), P.words.size() * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    S.ctx.s = sp;
    S.ctx.comb = D.comb;
    S.ctx.mlut = D.lut;
    S.ctx.clut = C.lut;
    S.ctx.perm = AdPermDev{(const PermSlotDev *)S.perm_slots.p, (const uint8_t *)S.perm_bytes.p,
                           (const uint32_t *)S.perm_words.p};
    S.ctx.zero = (T)0;
    S.ctx.max_scale = RateKind<T>::max_scale();
    S.err.alloc(64, s);
    HIPCHK(hipMemsetAsync(S.err.p, 0, 64, s));
    const T zero = S.ctx.zero, max_scale = S.ctx.max_scale;
    const T del_thr = (T)(so.deletion_threshold * (double)(max_scale - zero) + (double)zero);
    const double w4 = so.spawn_rate4, w2 = 1.0 - so.spawn_rate4;

    const bool files = o.pathname != nullptr;
    const std::string prefix = files ? o.pathname : "";
    const int steps = o.steps;
    const uint32_t ini = n_seeds ? seed_sum(seeds[0]) : 0u;
    std::vector<AdBookLayer> lay((size_t)steps);
    std::string stats_path;
    if (files) {
        stats_path = prefix + "stats.txt";
        if (!file_exists(stats_path)) {
            FILE *f = std::fopen(stats_path.c_str(), "a");
            if (!f) throw Error(EGTB_ERR_IO, "cannot write " + stats_path);
            std::fputs("layer,length,max success rate,speed,deletion_threshold,time\n", f);
            std::fclose(f);
        }
        if (res.started) {
            make_dir(prefix + std::to_string(steps - 1) + "b");
            make_dir(prefix + std::to_string(steps - 2) + "b");
        }
        ::unlink((prefix + std::to_string(steps - 2)).c_str());
    }
    S.describe(lay[steps - 1]);
    S.describe(lay[steps - 2]);
    bool started = false;
    for (int step = steps - 3; step >= 0; --step) {
        const std::string folder = prefix + std::to_string(step) + "b";
        if (files && is_dir(folder)) {
            ++res.solve_skipped;
            continue;
        }
        if (!started) {
            started = true;
            if (files && (step != steps - 3 || !res.started)) {
                S.read_layer(prefix + std::to_string(step + 1) + "b", lay[step + 1]);
                S.read_layer(prefix + std::to_string(step + 2) + "b", lay[step + 2]);
                S.index_layer(lay[step + 1]);
                S.index_layer(lay[step + 2]);
                S.describe(lay[step + 1]);
                S.describe(lay[step + 2]);
            }
        }
        const auto ts = std::chrono::steady_clock::now();
        // layer `step`: the forward pass's copy, else its raw file
        DBuf rawbuf;
        const uint64_t *raw = nullptr;
        uint64_t n = 0;
        if ((size_t)step < res.layers.size() && res.layers[step].p) {
            raw = (const uint64_t *)res.layers[step].p;
            n = res.n[step];
        } else if (files) {
            const std::vector<uint64_t> h = read_raw(prefix + std::to_string(step));
            n = h.size();
            rawbuf.alloc(n * 8 + 8, s);
            if (n) HIPCHK(hipMemcpyAsync(rawbuf.p, h.data(), n * 8, hipMemcpyHostToDevice, s));
            HIPCHK(hipStreamSynchronize(s));
            raw = (const uint64_t *)rawbuf.p;
        }
        const uint32_t bsum = (uint32_t)(2 * step) + ini;
        AdBookLayer &L0 = lay[step];
        S.bucketize(raw, n, bsum, pat.n_free_32k, L0);
        HIPCHK(hipMemsetAsync((uint64_t *)S.err.p + 1, 0, 8, s));
        for (int k = 0; k < kAdKeys; ++k)
            S.solve_bucket(k, L0.b[k], lay[step + 1].desc, lay[step + 2].desc, bsum, w2, w4);
        uint64_t ev[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(ev, S.err.p, sizeof ev, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const unsigned e = (unsigned)ev[0];
        if (e & kErrDerived) throw Error(EGTB_ERR_INTERNAL, "advanced solve: a board has more placements than its row");
        if (e & kErrColumn) throw Error(EGTB_ERR_INTERNAL, "advanced solve: a rate index outside its row");
        // length_count (:480-493): rates of the layer and the largest
        uint64_t len = 0;
        for (int k = 0; k < kAdKeys; ++k) len += L0.b[k].rows * L0.b[k].cols;
        T mx = zero;
        std::memcpy(&mx, &ev[1], sizeof(T));
        if (mx < zero) mx = zero;
        res.solve_values += len;
        if (so.deletion_threshold > 0.0) S.prune(lay[step + 2], del_thr);
        S.prune(L0, zero);
        S.index_layer(L0);  // layer `step` is searched by the next two steps
        S.describe(L0);
        HIPCHK(hipStreamSynchronize(s));
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
        if (files) {
            const double norm = (double)(mx - zero) / (double)(max_scale - zero);
            const double mbps = std::round((double)len / std::max(secs, 0.01) / 1e6 * 100.0) / 100.0;
            const std::string line = std::to_string(step) + "," + std::to_string(len) + "," + fmt_g(norm) + "," +
                                     fmt_g(mbps) + " mbps," + fmt_g(so.deletion_threshold) + "," + now_string() +
                                     "\n";
            FILE *f = std::fopen(stats_path.c_str(), "a");
            if (!f) throw Error(EGTB_ERR_IO, "cannot write " + stats_path);
            std::fputs(line.c_str(), f);
            std::fclose(f);
            if (so.deletion_threshold > 0.0) S.write_layer(prefix + std::to_string(step + 2) + "b", lay[step + 2]);
            S.write_layer(folder, L0);
            ::unlink((prefix + std::to_string(step)).c_str());
        }
        if (!so.keep_books) lay[step + 2] = AdBookLayer();
        if ((size_t)step < res.layers.size()) res.layers[step].release();
        progress_set((uint32_t)(steps - 2 - step), (uint32_t)(steps - 2));
    }
    HIPCHK(hipStreamSynchronize(s));
    res.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    res.rate_size = (int)sizeof(T);
    res.solved = true;
    if (so.keep_books) res.books = std::move(lay);
}

}  // namespace

void ad_solve(const uint64_t *seeds, size_t n_seeds, const egtb_ad_pattern &pat, const egtb_ad_options &o,
              const egtb_ad_solve_options &so, AdResult &res) {
    const std::string dt = so.success_rate_dtype ? so.success_rate_dtype : "float32";
    if (dt == "float32") ad_solve_t<float>(seeds, n_seeds, pat, o, so, res);
    else if (dt == "float64") ad_solve_t<double>(seeds, n_seeds, pat, o, so, res);
    else if (dt == "uint32") ad_solve_t<uint32_t>(seeds, n_seeds, pat, o, so, res);
    else if (dt == "uint64") ad_solve_t<uint64_t>(seeds, n_seeds, pat, o, so, res);
    else throw Error(EGTB_ERR_ARG, "advanced solve: success_rate_dtype must be float32, float64, uint32 or uint64");
}

}  // namespace egtb

struct egtb_ad_result : egtb::AdResult {};

extern "C" {

int egtb_ad_build(const uint64_t *seeds, size_t n_seeds, const egtb_ad_pattern *pattern,
                  const egtb_ad_options *options, const egtb_ad_solve_options *solve, egtb_ad_result **out) {
    if (out) *out = nullptr;
    if (!pattern || !options || !solve) {
        egtb::g_last_error = "pattern, options and solve options are required";
        return EGTB_ERR_ARG;
    }
    auto r = std::make_unique<egtb_ad_result>();
    const int rc = egtb::guarded_ad([&] { egtb::ad_solve(seeds, n_seeds, *pattern, *options, *solve, *r); });
    if (rc == EGTB_OK && out) *out = r.release();
    return rc;
}

int egtb_ad_result_solve_info(const egtb_ad_result *r, double *solve_ms, uint64_t *values, uint64_t *skipped_layers) {
    if (!r) {
        egtb::g_last_error = "result is NULL";
        return EGTB_ERR_ARG;
    }
    if (solve_ms) *solve_ms = r->solve_ms;
    if (values) *values = r->solve_values;
    if (skipped_layers) *skipped_layers = r->solve_skipped;
    return EGTB_OK;
}

int egtb_ad_result_book(const egtb_ad_result *r, int layer, int key, uint64_t *boards, void *rates, size_t cap_rows,
                        size_t *rows, int *cols) {
    return egtb::guarded_ad([&] {
        if (!r) throw egtb::Error(EGTB_ERR_ARG, "result is NULL");
        if (!r->solved || r->books.empty()) throw egtb::Error(EGTB_ERR_ARG, "books were not kept (solve.keep_books)");
        if (layer < 0 || layer >= (int)r->books.size()) throw egtb::Error(EGTB_ERR_ARG, "layer out of range");
        if (key < egtb::kAdKeyMin || key > egtb::kAdKeyMax) throw egtb::Error(EGTB_ERR_ARG, "key out of range");
        const egtb::AdBucket &B = r->books[layer].b[key - egtb::kAdKeyMin];
        if (rows) *rows = B.rows;
        if (cols) *cols = (int)B.cols;
        if ((boards || rates) && cap_rows < B.rows) throw egtb::Error(EGTB_ERR_ARG, "buffer too small");
        HIPCHK(hipSetDevice(r->dev));
        if (boards && B.rows) HIPCHK(hipMemcpy(boards, B.ind.p, B.rows * 8, hipMemcpyDeviceToHost));
        if (rates && B.rows && B.cols)
            HIPCHK(hipMemcpy(rates, B.rates.p, B.rows * B.cols * (size_t)r->rate_size, hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
