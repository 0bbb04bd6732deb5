// grouped.hpp - one kernel launch over the blocks of many problems (batched
// GEMMs of different shapes).
//
// A grouped launch runs the same per-block body as the single-problem kernel;
// block b belongs to the item p with start[p] <= b < start[p+1] and runs as
// block (b - start[p]) of that item.  Results are therefore identical to one
// launch per problem.  The host builds the items once; the engine copies them
// to device memory (d_items, d_start) so the launch can be replayed, also from
// a captured graph, without further host work.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

namespace oz2 {
namespace k {

struct GroupedLaunch {
    int kind = 0;                  // kernel variant (meaning private to the .hip file that built it)
    uint32_t n = 0;                // items
    uint32_t blocks = 0;           // total blocks (= start[n])
    size_t item_bytes = 0;
    std::vector<uint8_t> items;    // n packed items
    std::vector<uint32_t> start{0};
    std::vector<uint32_t> owner;   // optional: item of every block (one load instead of a search)
    const void* d_items = nullptr; // device copies, set by the owner before launching
    const uint32_t* d_start = nullptr;
    const uint32_t* d_owner = nullptr;

    template <typename T>
    void add(const T& item, uint32_t nblocks) {
        item_bytes = sizeof(T);
        const size_t at = items.size();
        items.resize(at + sizeof(T));
        std::memcpy(items.data() + at, &item, sizeof(T));
        blocks += nblocks;
        start.push_back(blocks);
        ++n;
    }
};

#if defined(__HIPCC__)
// item owning block b: the largest p with start[p] <= b (start[0] = 0)
__device__ __forceinline__ uint32_t owner_of(const uint32_t* __restrict__ start, uint32_t n, uint32_t b) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}
#endif

}  // namespace k
}  // namespace oz2
