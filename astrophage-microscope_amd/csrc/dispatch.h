// dispatch.h -- one kernel body, launched for one world or for a batch of
// independent worlds (an ensemble of chambers) in a single launch.
//
// Every per-tick kernel is a struct K with an argument type K::Args (the
// world's device view and scalars, by value) and a device body
// K::run(args, vgrid). `vgrid` is the body's own block index and block count:
//   * k_one<K>:  one world; vgrid = (blockIdx.x, gridDim.x);
//   * k_many<K>: chamber c owns a contiguous range of the launch's blocks,
//     found through a per-block map, and reads its arguments from args[c] in
//     device memory; vgrid is the block's index inside that range and the
//     range's size.
// Blocks of one chamber that share vgrid.block % NUM_XCD share an XCD, so the
// XCD-contiguous tile split of xcd_tiles holds per chamber.
//
// On the host, a batched launch goes through a Recorder: the per-chamber
// argument arrays and block prefixes are appended to one staging buffer, and
// the launch itself is deferred. The caller uploads the buffer once and then
// issues the recorded launches (typically under stream capture), so a whole
// ensemble tick is one graph whose kernels each cover every chamber.
#pragma once

#include <functional>
#include <map>
#include <string>
#include <span>
#include <vector>

#include "astro_dev.h"

namespace astro_dev {

// Block b of a batched launch belongs to chamber map[b].x, whose blocks start
// at map[b].y and number map[b].z.
template <class K>
__global__ __launch_bounds__(BLOCK) void k_one(typename K::Args a) {
    K::run(a, VGrid{static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x)});
}

// The body reads the chamber's arguments in place, through a restrict
// pointer: the block's chamber is uniform, so the fields load into scalar
// registers, and no store the body makes can alias them. (A local copy of the
// arguments would have its address taken by the body's lambdas and live in
// per-lane scratch memory.)
template <class K>
__global__ __launch_bounds__(BLOCK) void k_many(const typename K::Args* __restrict__ a,
                                                const int4* __restrict__ map) {
    const int4 m = map[blockIdx.x];
    K::run(a[m.x], VGrid{static_cast<int>(blockIdx.x) - m.y, m.z});
}

// The same wrappers with a register budget for at least W waves per SIMD.
template <class K, int W>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(W, 8))) void k_one_w(typename K::Args a) {
    K::run(a, VGrid{static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x)});
}

template <class K, int W>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(W, 8))) void k_many_w(
    const typename K::Args* __restrict__ a, const int4* __restrict__ map) {
    const int4 m = map[blockIdx.x];
    K::run(a[m.x], VGrid{static_cast<int>(blockIdx.x) - m.y, m.z});
}

// Host: batched launches of one ensemble step, recorded for later issue.
struct Recorder {
    struct Entry {
        const char* name;                                // stage name, for timing
        std::function<void(hipStream_t, const char*)> fn;  // issues the launch given the device arena
    };
    std::vector<char> staging;   // arguments and block maps, 256-byte aligned entries
    std::vector<Entry> launches;
    std::map<std::vector<int>, size_t> maps;   // block counts -> offset of their block map
    // Total blocks a batched launch of a grid-stride body is spread over
    // (0: one block per tile of each chamber's capacity, as a world alone),
    // by default and per stage name.
    int stride_blocks = 0;
    std::map<std::string, int> stage_blocks;
    // Lanes per x row of the float4 FTCS per stage name (0: widest that nx fills).
    std::map<std::string, int> ftcs_lanes;

    int blocks_for(const char* stage) const {
        const auto it = stage_blocks.find(stage);
        return it == stage_blocks.end() ? stride_blocks : it->second;
    }

    size_t put(const void* p, size_t n) {
        const size_t off = (staging.size() + 255) & ~static_cast<size_t>(255);
        staging.resize(off + n);
        std::copy(static_cast<const char*>(p), static_cast<const char*>(p) + n, staging.data() + off);
        return off;
    }
    // The block map of a launch with these per-chamber block counts.
    size_t block_map(const std::vector<int>& blocks) {
        const auto it = maps.find(blocks);
        if (it != maps.end()) return it->second;
        std::vector<int4> m;
        int first = 0;
        for (size_t c = 0; c < blocks.size(); ++c) {
            for (int b = 0; b < blocks[c]; ++b) m.push_back(int4{static_cast<int>(c), first, blocks[c], 0});
            first += blocks[c];
        }
        const size_t off = put(m.data(), sizeof(int4) * m.size());
        maps.emplace(blocks, off);
        return off;
    }
    void clear() {
        staging.clear();
        launches.clear();
        maps.clear();
    }
};

// Spreads a grid-stride launch over about `target` blocks in total, each
// chamber's share proportional to its block count: at least one block, and a
// multiple of NUM_XCD once it reaches NUM_XCD.
inline void spread_blocks(std::vector<int>& blocks, int target) {
    long long total = 0;
    for (int b : blocks) total += b;
    if (target <= 0 || total <= target) return;
    for (int& b : blocks) {
        int s = static_cast<int>((static_cast<long long>(b) * target + total - 1) / total);
        if (s < 1) s = 1;
        if (s >= NUM_XCD) s = ((s + NUM_XCD - 1) / NUM_XCD) * NUM_XCD;
        b = s < b ? s : b;
    }
}

}  // namespace astro_dev
