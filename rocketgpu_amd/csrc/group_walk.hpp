// Cooperative mesh-tree walk for the gfx950 step kernel: G lanes of one wavefront (aligned, G a power of two up to
// 32) share one arena and run its tick redundantly, with identical data, so every lane holds the same state.  Only
// the tree walks split the work.
//
// The stackless walk visits nodes in increasing index: from node i it moves to i + 1 when the node is a leaf or
// its box overlaps the query, else past the node's subtree (i + escape).  The group loads the window i .. i+G-1
// at once, one node per lane, with the overlap and leaf tests done per lane.  It then follows the walk through the
// window on ballot masks, so one memory latency covers up to G steps of the chain.  The leaves it visits compute
// their results in parallel, each on the lane that loaded it.  Those results are broadcast and committed on every
// lane in node order.  The walk visits the same leaves in the same order as SerialWalk and commits the same
// results, so the tick's bits are unchanged.
#pragma once

#include <hip/hip_runtime.h>

#include "core/rg_mesh.hpp"

namespace rg {

// Every 32-bit word of a trivially copyable record, read from lane `src`.
template <typename T>
__device__ inline T ShuffleRecord(const T& v, int src) {
    static_assert(sizeof(T) % 4 == 0, "records shuffle as 32-bit words");
    int in[sizeof(T) / 4];
    int out[sizeof(T) / 4];
    __builtin_memcpy(in, &v, sizeof(T));
#pragma unroll
    for (int w = 0; w < static_cast<int>(sizeof(T) / 4); ++w) out[w] = __shfl(in[w], src);
    T r;
    __builtin_memcpy(&r, out, sizeof(T));
    return r;
}

template <int G>
struct GroupWalk {
    static_assert(G >= 2 && G <= 32 && (G & (G - 1)) == 0, "group size must be a power of two in 2..32");
    // Bits of a landing code in 0..G.
    static constexpr int kLandBits = G >= 32 ? 6 : G >= 16 ? 5 : G >= 8 ? 4 : G >= 4 ? 3 : 2;

    // The group's G bits of a wave-wide ballot.
    __device__ static uint32_t Bits(unsigned long long ballot, int first) {
        return static_cast<uint32_t>((ballot >> first) & ((1ull << G) - 1ull));
    }

    // Returns the number of windows loaded (the walk's dependent memory rounds).
    template <typename Accept, typename SerialLeaf, typename Compute, typename Commit>
    __device__ static int Run(const core::MeshBvhNode* nodes, int begin, int end, const uint16_t* qmin,
                              const uint16_t* qmax, Accept&& accept, SerialLeaf&&, Compute&& compute,
                              Commit&& commit) {
        using Result = decltype(compute(int32_t{0}));
        const int lane = static_cast<int>(__lane_id());
        const int k = lane & (G - 1);
        const int first = lane - k;
        int base = begin;
        int windows = 0;
        while (base < end) {
            ++windows;
            const int idx = base + k;
            const bool valid = idx < end;
            core::MeshBvhNode node{};
            if (valid) node = nodes[idx];
            const bool overlap = valid && core::QuantizedOverlap(qmin, qmax, node.qmin, node.qmax);
            const bool leaf = valid && node.escape_or_tri >= 0;
            const bool take = overlap && leaf && accept(node.qmin, node.qmax);
            // Window-relative landing of each lane's node when the walk reaches it: the next node, or past its
            // subtree.  Capped at G + 1 ("beyond the window") and spread over kLandBits ballots, so the group can
            // follow the walk through the window on bitmasks alone.
            const int land = (overlap || leaf) ? k + 1 : k - node.escape_or_tri;
            const int code = (land < G + 1 ? land : G + 1) - 1;
            const uint32_t valid_m = Bits(__ballot(valid), first);
            const uint32_t next_m = Bits(__ballot(overlap || leaf), first);
            const uint32_t take_m = Bits(__ballot(take), first);
            uint32_t land_m[kLandBits];
#pragma unroll
            for (int b = 0; b < kLandBits; ++b) land_m[b] = Bits(__ballot((code >> b) & 1), first);
            uint32_t visit = 0;
            int pos = 0, far = -1;
            while (pos < G && ((valid_m >> pos) & 1u) != 0u) {
                // The run of nodes from pos that the walk steps through one by one, up to the first node it skips
                // past (or the end of the mesh's nodes).
                const uint64_t from = ~0ull << pos;
                const uint64_t stops = ((~static_cast<uint64_t>(next_m) | ~static_cast<uint64_t>(valid_m)) & from) |
                                       (1ull << G);
                const int q = __builtin_ctzll(stops);
                visit |= take_m & static_cast<uint32_t>(((1ull << q) - 1ull) & from);
                if (q >= G || ((valid_m >> q) & 1u) == 0u) {
                    pos = q;
                    break;
                }
                int c = 0;
#pragma unroll
                for (int b = 0; b < kLandBits; ++b) c |= static_cast<int>((land_m[b] >> q) & 1u) << b;
                if (c + 1 <= G) {
                    pos = c + 1;
                } else {
                    far = q;
                    break;
                }
            }
            const int far_land = __shfl(land, first + (far >= 0 ? far : 0));
            Result mine{};
            bool keep = false;
            if (((visit >> k) & 1u) != 0u) {
                mine = compute(node.escape_or_tri);
                keep = mine.keep();
            }
            // Only the results whose commit changes something travel between lanes, in node order.
            for (uint32_t m = Bits(__ballot(keep), first); m != 0u; m &= m - 1u) {
                commit(ShuffleRecord(mine, first + __builtin_ctz(m)));
            }
            base += far >= 0 ? far_land : pos;
        }
        return windows;
    }
};

}  // namespace rg
