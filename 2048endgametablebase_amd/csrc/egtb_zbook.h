// Compressed books in the reference's ".z" layout (trie index + xz blocks);
// host-side encode and decode.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace egtb {

struct ZBook {
    std::vector<uint64_t> keys;  // ascending, as in the book
    std::vector<uint8_t> rates;  // rate_size bytes each
};

// True when <pathname><i>.z/ (the compressed form of `book_path`, a path
// ending in ".book") exists.
bool zbook_exists(const std::string &book_path);
// The records of the compressed form of `book_path`.
ZBook zbook_read(const std::string &book_path, int rate_size);
// Writes the compressed form of a book of n sorted records (n > 0).
void zbook_write(const std::string &book_path, const uint64_t *keys, const void *rates, size_t n, int rate_size);

}  // namespace egtb
