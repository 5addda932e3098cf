// rg_layout.hpp -- offsets of a sequence of allocations and frees inside
// one device range (host code).
//
// Best fit among the freed ranges, neighbours merged, else the end of the
// layout; a freed range at the end shrinks it.  The same requests in the
// same order give the same offsets, so a dry run's peak sizes the range of
// a later run.
#pragma once

#include <stddef.h>

#include <iterator>
#include <map>
#include <unordered_map>

namespace rg {

struct OffsetLayout {
    std::map<size_t, size_t> free_ranges;  // offset -> bytes
    std::unordered_map<size_t, size_t> live;
    size_t top = 0, peak = 0;

    void reset() { free_ranges.clear(); live.clear(); top = peak = 0; }

    size_t take(size_t sz)
    {
        auto best = free_ranges.end();
        for (auto it = free_ranges.begin(); it != free_ranges.end(); ++it)
            if (it->second >= sz && (best == free_ranges.end() || it->second < best->second))
                best = it;
        size_t off;
        if (best != free_ranges.end()) {
            off = best->first;
            const size_t rest = best->second - sz;
            free_ranges.erase(best);
            if (rest) free_ranges[off + sz] = rest;
        } else {
            off = top;
            top += sz;
            if (top > peak) peak = top;
        }
        live[off] = sz;
        return off;
    }

    void give(size_t off)
    {
        auto lv = live.find(off);
        if (lv == live.end()) return;
        size_t sz = lv->second;
        live.erase(lv);
        auto nx = free_ranges.lower_bound(off);
        if (nx != free_ranges.end() && nx->first == off + sz) {
            sz += nx->second;
            nx = free_ranges.erase(nx);
        }
        if (nx != free_ranges.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) {
                off = pv->first;
                sz += pv->second;
                free_ranges.erase(pv);
            }
        }
        if (off + sz == top) top = off;  // the range ends the layout: shrink it
        else free_ranges[off] = sz;
    }
};

}  // namespace rg
