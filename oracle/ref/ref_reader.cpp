// Test infrastructure only: a C-ABI shim around the reference's own classic
// reader (ClassicBookReader::move_on_dic, native_core/src/ReaderRuntime.cpp:
// 1055-1068 -> evaluate_classic_result_candidates :739-840, and
// find_classic_value_native :1107-1133), compiled from the reference's
// sources by oracle/ref/Makefile into oracle/_ref/libegtb_ref.so.  Loaded by
// tests/golden/make_reader_golden.py to write reader fixtures, never by the
// product.
//
// egtb_ref_move_on_dic makes a fresh ClassicBookReader per call, so the
// symmetry ops are tried in the order 0, 0..7 (operation_sequence with
// last_operation_index_ = 0): the stateless order.  egtb_ref_reader_* keep one
// reader across calls, to show the reference's stateful order.

#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <utility>
#include <vector>

#include "FormationRuntime.h"
#include "ReaderRuntime.h"

namespace {

void set_err(char *err, size_t err_len, const char *msg) {
    if (err && err_len) {
        std::strncpy(err, msg, err_len - 1);
        err[err_len - 1] = '\0';
    }
}

PatternSpec make_spec(const char *name, const uint64_t *masks, size_t n_masks, int symm_mode) {
    PatternSpec spec;
    spec.name = name ? name : "";
    spec.pattern_masks.assign(masks, masks + n_masks);
    spec.symm_mode = symm_mode;
    return spec;
}

// Result layout: up to 4 entries; key i at keys[16*i] (NUL-terminated),
// kinds 0 none / 1 number / 2 string, numbers as doubles, strings at
// texts[16*i]; the dtype string goes to dtype_out (32 bytes).
int run_query(ClassicBookReader &reader, const int32_t *board16, const char *folder, const char *dtype,
              const char *pattern_full, int64_t nums_adjust, char *keys, int32_t *kinds, double *values,
              char *texts, int32_t *n_entries, char *dtype_out, char *err, size_t err_len) {
    try {
        std::vector<std::vector<int>> board(4, std::vector<int>(4));
        for (int r = 0; r < 4; ++r) {
            for (int c = 0; c < 4; ++c) {
                board[r][c] = board16[r * 4 + c];
            }
        }
        std::vector<std::pair<std::string, std::string>> paths;
        if (folder && folder[0]) {
            paths.emplace_back(folder, dtype);
        }
        ReaderMoveResult res = reader.move_on_dic(board, paths, pattern_full, nums_adjust);
        if (res.entries.size() > 4) {
            set_err(err, err_len, "more than 4 entries");
            return -3;
        }
        *n_entries = static_cast<int32_t>(res.entries.size());
        for (size_t i = 0; i < res.entries.size(); ++i) {
            const auto &e = res.entries[i];
            std::strncpy(keys + 16 * i, e.key.c_str(), 15);
            keys[16 * i + 15] = '\0';
            kinds[i] = e.kind == ReaderValueKind::NoneValue ? 0 : e.kind == ReaderValueKind::Numeric ? 1 : 2;
            values[i] = e.number;
            std::strncpy(texts + 16 * i, e.text.c_str(), 15);
            texts[16 * i + 15] = '\0';
        }
        std::strncpy(dtype_out, res.success_rate_dtype.c_str(), 31);
        dtype_out[31] = '\0';
        return 0;
    } catch (const std::exception &e) {
        set_err(err, err_len, e.what());
        return -1;
    } catch (...) {
        set_err(err, err_len, "unknown exception");
        return -2;
    }
}

}  // namespace

extern "C" int egtb_ref_move_on_dic(const int32_t *board16, const char *folder, const char *dtype,
                                    const char *pattern_full, int64_t nums_adjust, const char *name,
                                    const uint64_t *masks, size_t n_masks, int symm_mode, char *keys,
                                    int32_t *kinds, double *values, char *texts, int32_t *n_entries,
                                    char *dtype_out, char *err, size_t err_len) {
    ClassicBookReader reader(make_spec(name, masks, n_masks, symm_mode), false);
    return run_query(reader, board16, folder, dtype, pattern_full, nums_adjust, keys, kinds, values, texts,
                     n_entries, dtype_out, err, err_len);
}

extern "C" void *egtb_ref_reader_new(const char *name, const uint64_t *masks, size_t n_masks, int symm_mode) {
    try {
        return new ClassicBookReader(make_spec(name, masks, n_masks, symm_mode), false);
    } catch (...) {
        return nullptr;
    }
}

extern "C" int egtb_ref_reader_move_on_dic(void *handle, const int32_t *board16, const char *folder,
                                           const char *dtype, const char *pattern_full, int64_t nums_adjust,
                                           char *keys, int32_t *kinds, double *values, char *texts,
                                           int32_t *n_entries, char *dtype_out, char *err, size_t err_len) {
    return run_query(*static_cast<ClassicBookReader *>(handle), board16, folder, dtype, pattern_full,
                     nums_adjust, keys, kinds, values, texts, n_entries, dtype_out, err, err_len);
}

extern "C" void egtb_ref_reader_free(void *handle) { delete static_cast<ClassicBookReader *>(handle); }

extern "C" int egtb_ref_find_classic_value(const char *pathname, const char *filename, uint64_t key,
                                           const char *dtype, double *value, int32_t *found) {
    try {
        bool f = false;
        *value = find_classic_value_native(pathname, filename, key, dtype, f);
        *found = f ? 1 : 0;
        return 0;
    } catch (...) {
        return -1;
    }
}

// The reference's .z compression of one .book (trie_compress_progress_native,
// TrieCompression.cpp:538-551; the .book is left in place) and its per-probe
// search of a compressed layer (trie_decompress_search_native): 1 found with
// the normalised rate in *value, 0 when the search gives no value.
#include "TrieCompression.h"

extern "C" int egtb_ref_trie_compress(const char *book_path, const char *dtype) {
    try {
        return trie_compress_progress_native(book_path, dtype) ? 0 : -1;
    } catch (...) {
        return -2;
    }
}

extern "C" int egtb_ref_trie_search(const char *path_prefix, uint64_t board, const char *dtype, double *value) {
    try {
        const std::optional<double> v = trie_decompress_search_native(path_prefix, board, dtype);
        if (!v) return 0;
        *value = *v;
        return 1;
    } catch (...) {
        return -1;
    }
}
