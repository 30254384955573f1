// Compressed books: the reference's ".z" layout (TrieCompression.cpp:124-551,
// read by ReaderRuntime.cpp:318-476), written and read on the host.
//
// A book of sorted records {u64 board; rate} is stored in a directory
// <pathname><i>.z/ under the stem "<pattern>_<tile>_<i>." (the book's file
// name less "book"):
//
//   <stem>i   the first three trie levels, 5-byte nodes {u8 key; u32 next}:
//             a root {0, 1}, then the levels of byte 7, byte 6 and byte 5 of
//             the board, each led by a {0, 0} sentinel.  A node's `next` is
//             the position of its last child (levels 0 and 1 as positions in
//             this file, level 2 as an index into <stem>ii).
//   <stem>ii  the fourth level (byte 4), 3-byte nodes {u8 key; u16 next}: the
//             index of the node's last record relative to its block, 0 for
//             the node that ends a block.
//   <stem>s   12-byte entries {u32 node; u64 offset}: block k ends at level-4
//             node `node` and at byte `offset` of <stem>z; entry 0 is {0, 0}.
//   <stem>z   the blocks back to back, each an xz stream of packed
//             {u32 low 32 bits of the board; rate} records.  A block closes
//             after the first level-4 node that takes it past 32768 records.
//
// The blocks are xz (LZMA2, preset 1, CRC64) like the reference's; liblzma
// is loaded at run time, as the reference loads it.  Compression and
// decompression are host work; a decoded book goes to HBM like a .book.
#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <sys/stat.h>
#include <sys/types.h>

#include "egtb_runtime.h"
#include "egtb_zbook.h"

namespace egtb {

namespace {

// liblzma's stable C ABI (the few calls used here)
constexpr int kLzmaOk = 0;
constexpr int kLzmaBufError = 10;
constexpr int kLzmaCheckCrc64 = 4;
using EasyBufferEncode = int (*)(uint32_t preset, int check, const void *allocator, const uint8_t *in,
                                 size_t in_size, uint8_t *out, size_t *out_pos, size_t out_size);
using StreamBufferDecode = int (*)(uint64_t *memlimit, uint32_t flags, const void *allocator, const uint8_t *in,
                                   size_t *in_pos, size_t in_size, uint8_t *out, size_t *out_pos,
                                   size_t out_size);
using StreamBufferBound = size_t (*)(size_t size);

struct Lzma {
    EasyBufferEncode encode = nullptr;
    StreamBufferDecode decode = nullptr;
    StreamBufferBound bound = nullptr;
};

const Lzma &lzma() {
    static std::once_flag once;
    static Lzma api;
    std::call_once(once, [] {
        for (const char *name : {"liblzma.so.5", "liblzma.so"}) {
            void *h = dlopen(name, RTLD_NOW | RTLD_LOCAL);
            if (!h) continue;
            api.encode = (EasyBufferEncode)dlsym(h, "lzma_easy_buffer_encode");
            api.decode = (StreamBufferDecode)dlsym(h, "lzma_stream_buffer_decode");
            api.bound = (StreamBufferBound)dlsym(h, "lzma_stream_buffer_bound");
            if (api.encode && api.decode && api.bound) return;
            api = Lzma{};
        }
    });
    if (!api.encode) throw Error(EGTB_ERR_IO, "liblzma (liblzma.so.5) could not be loaded for compressed books");
    return api;
}

#pragma pack(push, 1)
struct Node32 {
    uint8_t key;
    uint32_t next;
};
struct Node16 {
    uint8_t key;
    uint16_t next;
};
struct Segment {
    uint32_t node;
    uint64_t offset;
};
#pragma pack(pop)
static_assert(sizeof(Node32) == 5 && sizeof(Node16) == 3 && sizeof(Segment) == 12, "packed .z layouts");

constexpr size_t kBlockRecords = 32768;

std::vector<uint8_t> read_all(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error(EGTB_ERR_IO, "cannot open " + path);
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<uint8_t> out(sz > 0 ? (size_t)sz : 0);
    const size_t got = out.empty() ? 0 : std::fread(out.data(), 1, out.size(), f);
    std::fclose(f);
    if (got != out.size()) throw Error(EGTB_ERR_IO, "short read of " + path);
    return out;
}

template <typename N>
std::vector<N> read_nodes(const std::string &path) {
    const std::vector<uint8_t> b = read_all(path);
    if (b.size() % sizeof(N)) throw Error(EGTB_ERR_IO, path + ": size is not a multiple of its node size");
    std::vector<N> v(b.size() / sizeof(N));
    if (!v.empty()) std::memcpy(v.data(), b.data(), b.size());
    return v;
}

void write_all(const std::string &path, const void *data, size_t bytes) { write_file_atomic(path, data, bytes); }

std::vector<uint8_t> xz_decode(const uint8_t *in, size_t n, size_t expect) {
    const Lzma &api = lzma();
    size_t cap = std::max<size_t>(expect ? expect : n * 8, 4096);
    for (int attempt = 0; attempt < 12; ++attempt) {
        std::vector<uint8_t> out(cap);
        uint64_t memlimit = UINT64_MAX;
        size_t in_pos = 0, out_pos = 0;
        const int rc = api.decode(&memlimit, 0, nullptr, in, &in_pos, n, out.data(), &out_pos, out.size());
        if (rc == kLzmaOk) {
            out.resize(out_pos);
            return out;
        }
        if (rc != kLzmaBufError) throw Error(EGTB_ERR_IO, "xz block does not decode (liblzma " + std::to_string(rc) + ")");
        cap *= 2;
    }
    throw Error(EGTB_ERR_IO, "xz block too large");
}

std::vector<uint8_t> xz_encode(const uint8_t *in, size_t n) {
    const Lzma &api = lzma();
    std::vector<uint8_t> out(api.bound(n));
    size_t pos = 0;
    const int rc = api.encode(1, kLzmaCheckCrc64, nullptr, in, n, out.data(), &pos, out.size());
    if (rc != kLzmaOk) throw Error(EGTB_ERR_IO, "xz compression failed (liblzma " + std::to_string(rc) + ")");
    out.resize(pos);
    return out;
}

// Runs f(k) for k in [0, n) on up to `threads` host threads.
template <typename F>
void parallel_for(size_t n, int threads, F &&f) {
    threads = std::max(1, std::min<int>(threads, (int)n));
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    std::exception_ptr err;
    std::mutex emu;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            try {
                for (size_t k; (k = next.fetch_add(1)) < n;) f(k);
            } catch (...) {
                std::lock_guard<std::mutex> lk(emu);
                if (!err) err = std::current_exception();
                next.store(n);
            }
        });
    for (auto &th : pool) th.join();
    if (err) std::rethrow_exception(err);
}

int host_threads() {
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(hc ? hc : 4u, 16u));
}

}  // namespace

bool zbook_exists(const std::string &book_path) {
    struct stat st;
    const std::string dir = book_path.substr(0, book_path.size() - 5) + ".z";
    return stat(dir.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

static std::string zbook_prefix(const std::string &book_path) {
    // <pathname><i>.book -> <pathname><i>.z/<stem>, the stem being the file
    // name less its last four characters, so it keeps the dot:
    // "442_256_95.book" -> "442_256_95." (TrieCompression.cpp:320-323,
    // ReaderRuntime.cpp:494-495)
    const std::string base = book_path.substr(0, book_path.size() - 5);
    const size_t slash = book_path.find_last_of('/');
    const std::string name = slash == std::string::npos ? book_path : book_path.substr(slash + 1);
    return base + ".z/" + name.substr(0, name.size() - 4);
}

ZBook zbook_read(const std::string &book_path, int rate_size) {
    const std::string pre = zbook_prefix(book_path);
    const std::vector<Node32> ind = read_nodes<Node32>(pre + "i");
    const std::vector<Node16> ii = read_nodes<Node16>(pre + "ii");
    const std::vector<Segment> seg = read_nodes<Segment>(pre + "s");
    const std::vector<uint8_t> z = read_all(pre + "z");
    const size_t rec = 4 + (size_t)rate_size;
    if (ind.size() < 2 || seg.size() < 2 || ii.size() < 2)
        throw Error(EGTB_ERR_IO, pre + ": trie index too short");
    // the three flattened levels: root, level 0 (byte 7), level 1 (byte 6), level 2 (byte 5)
    const size_t n0 = ind[1].next - 1;
    const size_t s1 = 1 + n0;
    if (s1 >= ind.size()) throw Error(EGTB_ERR_IO, pre + ": malformed trie index");
    const size_t n1 = ind[s1].next - s1;
    const size_t s2 = s1 + n1;
    if (s2 > ind.size()) throw Error(EGTB_ERR_IO, pre + ": malformed trie index");
    // the board's bytes 7..5 of every level-4 node (index into ii)
    std::vector<uint32_t> high(ii.size(), 0);
    for (size_t a = 2; a <= n0; ++a) {  // level 0 nodes at positions 2..n0 (position 1 is its sentinel)
        const uint32_t b7 = ind[a].key;
        for (size_t c = ind[a - 1].next + 1; c <= ind[a].next; ++c) {
            const uint32_t b6 = ind[c].key;
            for (size_t d = ind[c - 1].next + 1; d <= ind[c].next; ++d) {
                const uint32_t b5 = ind[d].key;
                for (size_t e = ind[d - 1].next + 1; e <= ind[d].next && e < ii.size(); ++e)
                    high[e] = (b7 << 16) | (b6 << 8) | b5;
            }
        }
    }
    // blocks: decompress in parallel, then assign each record its node
    const size_t nblocks = seg.size() - 1;
    std::vector<std::vector<uint8_t>> blocks(nblocks);
    parallel_for(nblocks, host_threads(), [&](size_t k) {
        const uint64_t a = seg[k].offset, b = seg[k + 1].offset;
        if (b < a || b > z.size()) throw Error(EGTB_ERR_IO, pre + ": block offsets outside the z file");
        blocks[k] = xz_decode(z.data() + a, (size_t)(b - a), (kBlockRecords + 64) * rec);
        if (blocks[k].size() % rec) throw Error(EGTB_ERR_IO, pre + ": block of partial records");
    });
    size_t total = 0;
    for (const auto &b : blocks) total += b.size() / rec;
    ZBook out;
    out.keys.resize(total);
    out.rates.resize(total * (size_t)rate_size);
    size_t base = 0;
    for (size_t k = 0; k < nblocks; ++k) {
        const size_t m = blocks[k].size() / rec;
        const size_t first = seg[k].node + 1, last = seg[k + 1].node;  // level-4 nodes of this block
        size_t r = 0;
        for (size_t e = first; e <= last && e < ii.size(); ++e) {
            const size_t end = e == last ? m : (size_t)ii[e].next + 1;  // records [r, end)
            if (end > m || end < r) throw Error(EGTB_ERR_IO, pre + ": trie node outside its block");
            const uint64_t hi = ((uint64_t)high[e] << 40) | ((uint64_t)ii[e].key << 32);
            for (; r < end; ++r) {
                uint32_t lo;
                std::memcpy(&lo, blocks[k].data() + r * rec, 4);
                out.keys[base + r] = hi | lo;
                std::memcpy(out.rates.data() + (base + r) * (size_t)rate_size, blocks[k].data() + r * rec + 4,
                            (size_t)rate_size);
            }
        }
        if (r != m) throw Error(EGTB_ERR_IO, pre + ": block records without a trie node");
        base += m;
    }
    return out;
}

void zbook_write(const std::string &book_path, const uint64_t *keys, const void *rates, size_t n, int rate_size) {
    if (n == 0) throw Error(EGTB_ERR_ARG, "an empty book is not compressed");
    const size_t rec = 4 + (size_t)rate_size;
    // the trie over bytes 7, 6, 5, 4 (build_trie_index's grouping): each
    // level's node keeps its key and the index of its last child
    std::vector<Node32> l0{{0, 0}}, l1{{0, 0}}, l2{{0, 0}};
    std::vector<uint64_t> l3_last{0};
    std::vector<uint8_t> l3_key{0};
    auto byte = [](uint64_t b, int q) { return (uint8_t)((b >> (8 * q)) & 0xFF); };
    for (size_t j = 1; j <= n; ++j) {
        const bool end = j == n;
        const uint64_t prev = keys[j - 1];
        const int change = end                                     ? 7
                           : byte(keys[j], 7) != byte(prev, 7) ? 7
                           : byte(keys[j], 6) != byte(prev, 6) ? 6
                           : byte(keys[j], 5) != byte(prev, 5) ? 5
                           : byte(keys[j], 4) != byte(prev, 4) ? 4
                                                                  : 0;
        if (change >= 4) {
            l3_key.push_back(byte(prev, 4));
            l3_last.push_back(j - 1);
        }
        if (change >= 5) l2.push_back({byte(prev, 5), (uint32_t)(l3_key.size() - 1)});
        if (change >= 6) l1.push_back({byte(prev, 6), (uint32_t)(l2.size() - 1)});
        if (change >= 7) l0.push_back({byte(prev, 7), (uint32_t)(l1.size() - 1)});
    }
    // blocks: close after the first level-4 node past kBlockRecords records
    struct Blk {
        size_t start, end;
        uint32_t node;
    };
    std::vector<Blk> blks;
    std::vector<Node16> ii(l3_key.size());
    ii[0] = {0, 0};
    size_t start = 0;
    for (size_t e = 1; e < l3_key.size(); ++e) {
        const size_t end = (size_t)l3_last[e] + 1;
        ii[e].key = l3_key[e];
        if (e + 1 == l3_key.size() || end - start > kBlockRecords) {
            const size_t stop = e + 1 == l3_key.size() ? n : end;
            blks.push_back({start, stop, (uint32_t)e});
            ii[e].next = 0;
            start = stop;
        } else {
            ii[e].next = (uint16_t)(l3_last[e] - start);
        }
    }
    std::vector<std::vector<uint8_t>> z(blks.size());
    parallel_for(blks.size(), host_threads(), [&](size_t k) {
        const Blk &b = blks[k];
        std::vector<uint8_t> raw((b.end - b.start) * rec);
        for (size_t r = b.start; r < b.end; ++r) {
            const uint32_t lo = (uint32_t)(keys[r] & 0xFFFFFFFFull);
            std::memcpy(raw.data() + (r - b.start) * rec, &lo, 4);
            std::memcpy(raw.data() + (r - b.start) * rec + 4, (const char *)rates + r * (size_t)rate_size,
                        (size_t)rate_size);
        }
        z[k] = xz_encode(raw.data(), raw.size());
    });
    std::vector<Segment> seg{{0, 0}};
    std::vector<uint8_t> zfile;
    for (size_t k = 0; k < blks.size(); ++k) {
        zfile.insert(zfile.end(), z[k].begin(), z[k].end());
        seg.push_back({blks[k].node, (uint64_t)zfile.size()});
    }
    // flatten levels 0..2 behind a root, their `next` as positions
    std::vector<Node32> ind{{0, 1}};
    const uint32_t off0 = 1 + (uint32_t)l0.size(), off1 = off0 + (uint32_t)l1.size();
    for (Node32 v : l0) ind.push_back({v.key, v.next + off0});
    for (Node32 v : l1) ind.push_back({v.key, v.next + off1});
    for (Node32 v : l2) ind.push_back(v);
    ind[1].next = off0;  // level 0's sentinel points at the start of level 1
    ind[off0].next = off1;
    const std::string base = book_path.substr(0, book_path.size() - 5);
    const std::string dir = base + ".z";
    mkdir(dir.c_str(), 0777);
    const std::string pre = zbook_prefix(book_path);
    write_all(pre + "z", zfile.data(), zfile.size());
    write_all(pre + "ii", ii.data(), ii.size() * sizeof(Node16));
    write_all(pre + "i", ind.data(), ind.size() * sizeof(Node32));
    write_all(pre + "s", seg.data(), seg.size() * sizeof(Segment));
}

}  // namespace egtb
