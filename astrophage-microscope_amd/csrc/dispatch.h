// dispatch.h -- one kernel body, launched for one world or for a batch of
// independent worlds (an ensemble of chambers) in a single launch.
//
// Every per-tick kernel is a struct K with an argument type K::Args (the
// world's device view and scalars, by value) and a device body
// K::run(args, vgrid). `vgrid` is the body's own block index and block count:
//   * k_one<K>:  one world; vgrid = (blockIdx.x, gridDim.x);
//   * k_many<K>: chamber c owns blocks [first[c], first[c + 1]) of the launch
//     and reads its arguments from args[c] in device memory; vgrid is the
//     block's index inside that range and the range's size.
// Each chamber's block count is a multiple of NUM_XCD, so a block's XCD
// (blockIdx.x % NUM_XCD) equals its virtual block's, and the XCD-contiguous
// tile split of xcd_tiles holds per chamber.
//
// On the host, a batched launch goes through a Recorder: the per-chamber
// argument arrays and block prefixes are appended to one staging buffer, and
// the launch itself is deferred. The caller uploads the buffer once and then
// issues the recorded launches (typically under stream capture), so a whole
// ensemble tick is one graph whose kernels each cover every chamber.
#pragma once

#include <functional>
#include <span>
#include <vector>

#include "astro_dev.h"

namespace astro_dev {

__device__ __forceinline__ int chamber_of(const int* first, int nch) {
    const int b = static_cast<int>(blockIdx.x);
    int lo = 0, hi = nch;   // largest c with first[c] <= b
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <class K>
__global__ __launch_bounds__(BLOCK) void k_one(typename K::Args a) {
    K::run(a, VGrid{static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x)});
}

// The chamber's arguments are copied out of device memory first: read through
// the pointer, every field would be re-read after each store the body makes.
template <class K>
__global__ __launch_bounds__(BLOCK) void k_many(const typename K::Args* a, const int* first, int nch) {
    const int c = chamber_of(first, nch);
    const typename K::Args args = a[c];
    K::run(args, VGrid{static_cast<int>(blockIdx.x) - first[c], first[c + 1] - first[c]});
}

// The same wrappers for bodies that need three waves per SIMD of registers.
template <class K>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_one_w3(typename K::Args a) {
    K::run(a, VGrid{static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x)});
}

template <class K>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_many_w3(
    const typename K::Args* a, const int* first, int nch) {
    const int c = chamber_of(first, nch);
    const typename K::Args args = a[c];
    K::run(args, VGrid{static_cast<int>(blockIdx.x) - first[c], first[c + 1] - first[c]});
}

// Host: batched launches of one ensemble step, recorded for later issue.
struct Recorder {
    struct Entry {
        const char* name;                                // stage name, for timing
        std::function<void(hipStream_t, const char*)> fn;  // issues the launch given the device arena
    };
    std::vector<char> staging;   // arguments and block prefixes, 256-byte aligned entries
    std::vector<Entry> launches;

    size_t put(const void* p, size_t n) {
        const size_t off = (staging.size() + 255) & ~static_cast<size_t>(255);
        staging.resize(off + n);
        std::copy(static_cast<const char*>(p), static_cast<const char*>(p) + n, staging.data() + off);
        return off;
    }
    void clear() {
        staging.clear();
        launches.clear();
    }
};

}  // namespace astro_dev
