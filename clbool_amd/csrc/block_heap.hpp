// block_heap.hpp -- best-fit sub-allocation inside segments, with splitting
// and coalescing; the host-side bookkeeping of the device buffer pool
// (runtime.hip), kept free of device calls so that it builds and is tested
// with the host compiler alone.
//
// A segment is one contiguous range handed in by add_segment().  Its blocks
// tile it in address order (prev/next); take(cap) returns the smallest free
// block of at least cap bytes, split when the rest is at least min_split;
// give_back() marks a block free and merges it with free neighbours, so a
// segment whose blocks are all free is a single free block again, which
// collect_free_segments() hands out for release.
#pragma once

#include <cstdint>
#include <map>
#include <set>
#include <utility>
#include <vector>

namespace clb {

class BlockHeap {
  public:
    struct Block {
        char *ptr;
        std::uint64_t size;
        bool free;
        Block *prev, *next;  // address neighbours inside the segment
    };

    explicit BlockHeap(std::uint64_t min_split) : min_split_(min_split) {}
    BlockHeap(const BlockHeap &) = delete;
    BlockHeap &operator=(const BlockHeap &) = delete;
    ~BlockHeap() {
        for (auto &kv : free_) delete kv.second;
        for (auto &kv : live_) delete kv.second;
    }

    // Smallest free block of at least cap bytes, split; nullptr if none.
    Block *take(std::uint64_t cap) {
        auto it = free_.lower_bound({cap, nullptr});
        if (it == free_.end()) return nullptr;
        Block *b = it->second;
        free_.erase(it);
        free_bytes_ -= b->size;
        if (b->size - cap >= min_split_) {
            Block *r = new Block{b->ptr + cap, b->size - cap, true, b, b->next};
            if (b->next) b->next->prev = r;
            b->next = r;
            b->size = cap;
            free_.insert({r->size, r});
            free_bytes_ += r->size;
        }
        b->free = false;
        live_[b->ptr] = b;
        live_bytes_ += b->size;
        return b;
    }

    void add_segment(char *base, std::uint64_t size) {
        Block *b = new Block{base, size, true, nullptr, nullptr};
        free_.insert({size, b});
        free_bytes_ += size;
        seg_bytes_ += size;
    }

    // Frees the block that starts at p; false if no live block starts there.
    bool give_back(char *p) {
        auto it = live_.find(p);
        if (it == live_.end()) return false;
        Block *b = it->second;
        live_.erase(it);
        live_bytes_ -= b->size;
        b->free = true;
        if (b->prev && b->prev->free) {
            Block *a = b->prev;
            unlist(a);
            a->size += b->size;
            a->next = b->next;
            if (b->next) b->next->prev = a;
            delete b;
            b = a;
        }
        if (b->next && b->next->free) {
            Block *c = b->next;
            unlist(c);
            b->size += c->size;
            b->next = c->next;
            if (c->next) c->next->prev = b;
            delete c;
        }
        free_.insert({b->size, b});
        free_bytes_ += b->size;
        return true;
    }

    // Removes every segment that is one free block and appends (base, size).
    void collect_free_segments(std::vector<std::pair<char *, std::uint64_t>> &out) {
        for (auto it = free_.begin(); it != free_.end();) {
            Block *b = it->second;
            if (!b->prev && !b->next) {
                out.push_back({b->ptr, b->size});
                free_bytes_ -= b->size;
                seg_bytes_ -= b->size;
                delete b;
                it = free_.erase(it);
            } else {
                ++it;
            }
        }
    }

    std::uint64_t free_bytes() const { return free_bytes_; }
    std::uint64_t live_bytes() const { return live_bytes_; }
    std::uint64_t segment_bytes() const { return seg_bytes_; }
    std::uint64_t largest_free() const { return free_.empty() ? 0 : free_.rbegin()->first; }
    std::size_t live_count() const { return live_.size(); }
    // Live blocks in address order (for checks).
    const std::map<char *, Block *> &live() const { return live_; }

  private:
    void unlist(Block *b) {
        free_.erase({b->size, b});
        free_bytes_ -= b->size;
    }
    std::uint64_t min_split_;
    std::set<std::pair<std::uint64_t, Block *>> free_;  // (size, block)
    std::map<char *, Block *> live_;
    std::uint64_t free_bytes_ = 0, live_bytes_ = 0, seg_bytes_ = 0;
};

}  // namespace clb
