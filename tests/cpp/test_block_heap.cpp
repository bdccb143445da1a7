// Randomised check of the device pool's block bookkeeping (BlockHeap):
// taken blocks lie inside their segment, never overlap, and are at least the
// size asked for; free + live bytes always equal the segments' bytes; a
// released block merges with free neighbours, so once everything is released
// every segment is one free block and is collected whole.  Addresses are
// synthetic (never dereferenced).  Exit status 0 on success.
#include "../../clbool_amd/csrc/block_heap.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

using clb::BlockHeap;
using u64 = std::uint64_t;

#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            return 1;                                                        \
        }                                                                    \
    } while (0)

struct Seg {
    u64 base, size;
};

static int run(u64 unit, u64 min_split, u64 seg_min, unsigned seed, int ops) {
    BlockHeap h(min_split);
    std::mt19937_64 rng(seed);
    std::vector<Seg> segs;
    std::vector<std::pair<char *, u64>> held;  // (ptr, size asked)
    u64 next_base = 1ull << 40;
    u64 total = 0;
    auto seg_of = [&](u64 a) -> const Seg * {
        for (const Seg &s : segs)
            if (a >= s.base && a < s.base + s.size) return &s;
        return nullptr;
    };
    for (int op = 0; op < ops; ++op) {
        const bool take = held.empty() || (rng() % 100) < 55;
        if (take) {
            // sizes over four orders of magnitude, in units
            const u64 units = 1 + (rng() % 4 == 0 ? rng() % 4096 : rng() % 16);
            const u64 cap = units * unit;
            BlockHeap::Block *b = h.take(cap);
            if (!b) {
                const u64 sz = cap > seg_min ? cap : seg_min;
                segs.push_back({next_base, sz});
                h.add_segment(reinterpret_cast<char *>(next_base), sz);
                next_base += sz + (1ull << 30);  // segments never touch
                total += sz;
                b = h.take(cap);
                CHECK(b != nullptr);
            }
            CHECK(b->size >= cap);
            CHECK(b->size - cap < min_split || b->size == cap);
            const u64 a = reinterpret_cast<u64>(b->ptr);
            const Seg *s = seg_of(a);
            CHECK(s != nullptr);
            CHECK(a + b->size <= s->base + s->size);
            held.push_back({b->ptr, cap});
        } else {
            const size_t i = rng() % held.size();
            CHECK(h.give_back(held[i].first));
            CHECK(!h.give_back(held[i].first));  // a second release is refused
            held[i] = held.back();
            held.pop_back();
        }
        CHECK(h.free_bytes() + h.live_bytes() == total);
        CHECK(h.segment_bytes() == total);
        CHECK(h.live_count() == held.size());
        if (op % 97 == 0) {  // live blocks are disjoint, in address order
            u64 end = 0;
            for (auto &kv : h.live()) {
                const u64 a = reinterpret_cast<u64>(kv.first);
                CHECK(a >= end);
                end = a + kv.second->size;
            }
        }
    }
    // a free segment is collected only when entirely free
    std::vector<std::pair<char *, u64>> out;
    h.collect_free_segments(out);
    for (auto &f : out) {
        const Seg *s = seg_of(reinterpret_cast<u64>(f.first));
        CHECK(s && s->base == reinterpret_cast<u64>(f.first) && s->size == f.second);
        for (auto &p : held) {
            const u64 a = reinterpret_cast<u64>(p.first);
            CHECK(a < s->base || a >= s->base + s->size);
        }
        total -= f.second;
    }
    CHECK(h.segment_bytes() == total);
    for (auto &p : held) CHECK(h.give_back(p.first));
    CHECK(h.live_bytes() == 0);
    CHECK(h.free_bytes() == total);
    out.clear();
    h.collect_free_segments(out);
    u64 got = 0;
    for (auto &f : out) got += f.second;
    CHECK(got == total);
    CHECK(h.free_bytes() == 0 && h.segment_bytes() == 0 && h.largest_free() == 0);
    // coalescing: a segment split into pieces and released in any order is
    // one block again, which serves a request of the whole segment (heaps
    // that split off remainders of one unit)
    if (min_split > unit) return 0;
    BlockHeap g(min_split);
    const u64 seg = 64 * unit;
    g.add_segment(reinterpret_cast<char *>(next_base), seg);
    std::vector<char *> pieces;
    for (int k = 0; k < 64; ++k) {
        BlockHeap::Block *b = g.take(unit);
        CHECK(b && b->ptr == reinterpret_cast<char *>(next_base + k * unit));
        pieces.push_back(b->ptr);
    }
    CHECK(g.take(unit) == nullptr);
    std::shuffle(pieces.begin(), pieces.end(), rng);
    for (char *p : pieces) CHECK(g.give_back(p));
    CHECK(g.largest_free() == seg);
    BlockHeap::Block *w = g.take(seg);
    CHECK(w && w->size == seg);
    return 0;
}

int main() {
    int rc = 0;
    // the pool's two heaps: 256-byte small blocks in 2 MiB segments, 2 MiB
    // large blocks in segments of at least 1 GiB
    rc |= run(256, 256, 2ull << 20, 1, 200000);
    rc |= run(2ull << 20, 2ull << 20, 1ull << 30, 2, 200000);
    // a coarse split threshold keeps small remainders inside the block
    rc |= run(256, 4096, 1ull << 20, 3, 100000);
    if (rc == 0) std::printf("block heap: ok\n");
    return rc;
}
