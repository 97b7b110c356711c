// scan.h -- exclusive prefix sums with no inter-block waiting, gfx950.
//
// Replaces the reference's cub::DeviceScan::ExclusiveSum call sites
// (src/sim/cell_store.cu:422, src/sim/lifecycle.cu:269) and the cell-range
// half of the hash build. An exclusive scan's output is uniquely defined, so
// any correct scan is bit-identical to CUB's.
//
// Every array scanned here is produced by an earlier kernel of the same tick
// (bucket populations, divide flags, drop flags), and that producer also adds
// each item into a per-tile total, `tile_sum[i >> SCAN_TILE_SHIFT]`. The scan
// is then a single kernel with one block per 4096-item tile: the block sums
// the totals of the tiles before it (a few thousand ints at most, L2-resident),
// scans its own tile in registers and LDS (256 threads x 16 items, 16-byte
// vector loads and stores) and writes. No block ever waits for another, so
// there is no spin, no dependence on dispatch order or on when another XCD's
// stores become visible, and the kernel can sit in a captured graph. A tile
// whose own total disagrees with the producer's raises ERR_SCAN in the
// device error flags, which the host checks.
#pragma once

#include "astro_dev.h"

namespace astro_dev {

constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE_SHIFT = 12;
constexpr int SCAN_TILE = 1 << SCAN_TILE_SHIFT;   // BLOCK * SCAN_ITEMS
static_assert(SCAN_TILE == BLOCK * SCAN_ITEMS, "scan tile shape");

__host__ __device__ __forceinline__ int scan_tiles_for(int n) { return n <= 0 ? 0 : (n + SCAN_TILE - 1) / SCAN_TILE; }

__device__ __forceinline__ int32_t wave_inclusive_sum(int32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

__device__ __forceinline__ int32_t wave_sum(int32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Adds, for every active lane, 1 to arr[key]: one atomic per distinct key in
// the wave. Lanes of a wave walking sorted data mostly share a key, so this is
// usually one atomic per wave. Every lane of the wave must call it.
__device__ __forceinline__ void wave_count_by_key(int32_t* arr, uint32_t key, bool active) {
    unsigned long long pending = __ballot(active);
    const int lane = threadIdx.x & 63;
    while (pending) {
        const int leader = __ffsll(static_cast<long long>(pending)) - 1;
        const uint32_t k = static_cast<uint32_t>(__shfl(static_cast<int>(key), leader, 64));
        const unsigned long long same = __ballot(active && key == k) & pending;
        if (lane == leader) atomicAdd(&arr[k], static_cast<int32_t>(__popcll(same)));
        pending &= ~same;
    }
}

// Vector helpers for a plain int32 array.
__device__ __forceinline__ int4 load4_i32(const int32_t* a, int i, int n) {
    if (i + 3 < n) return *reinterpret_cast<const int4*>(a + i);
    int4 v{0, 0, 0, 0};
    if (i < n) v.x = a[i];
    if (i + 1 < n) v.y = a[i + 1];
    if (i + 2 < n) v.z = a[i + 2];
    return v;
}
__device__ __forceinline__ void store4_i32(int32_t* a, int i, int n, int4 v) {
    if (i + 3 < n) {
        *reinterpret_cast<int4*>(a + i) = v;
        return;
    }
    if (i < n) a[i] = v.x;
    if (i + 1 < n) a[i + 1] = v.y;
    if (i + 2 < n) a[i + 2] = v.z;
}

// Block-wide sum; every thread gets the result. s_wave holds BLOCK/64 ints.
__device__ __forceinline__ int32_t block_sum(int32_t v, int32_t* s_wave) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = v;
    __syncthreads();
    int32_t t = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) t += s_wave[w];
    return t;
}

// Exclusive scan of tile `tile` of the n items produced by load4(i) (items
// i..i+3, i a multiple of 4, zero past n). store4(i, prefixes, items) receives
// the exclusive prefixes of items i..i+3 and the items themselves. tile_sum[t]
// must hold the sum of tile t's items. The last tile writes the grand total to
// *total_out. Called by every thread of one block.
template <class Load4, class Store4>
__device__ void scan_tile(int n, int tile, Load4 load4, Store4 store4, const int32_t* tile_sum, int32_t* total_out,
                          uint32_t* error_flags) {
    __shared__ int32_t s_wave[BLOCK / 64];
    const int num_tiles = scan_tiles_for(n);
    if (tile >= num_tiles) return;
    // The tile's own items are loaded first: they do not depend on the
    // prefix, so their latency overlaps the sum of the earlier tiles' totals.
    const int first = tile * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int4 v[SCAN_ITEMS / 4];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS / 4; ++k) v[k] = load4(first + 4 * k);
    int32_t before = 0;
    for (int t = threadIdx.x; t < tile; t += BLOCK) before += tile_sum[t];
    const int32_t prefix = block_sum(before, s_wave);

    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS / 4; ++k) sum += v[k].x + v[k].y + v[k].z + v[k].w;
    const int32_t incl = wave_inclusive_sum(sum);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    int32_t wave_off = 0, tile_total = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) {
        if (w < wave) wave_off += s_wave[w];
        tile_total += s_wave[w];
    }
    if (threadIdx.x == 0) {
        if (tile_total != tile_sum[tile]) atomicOr(error_flags, ERR_SCAN);
        if (tile == num_tiles - 1 && total_out) *total_out = prefix + tile_total;
    }
    int32_t run = prefix + wave_off + incl - sum;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS / 4; ++k) {
        int4 o;
        o.x = run; run += v[k].x;
        o.y = run; run += v[k].y;
        o.z = run; run += v[k].z;
        o.w = run; run += v[k].w;
        store4(first + 4 * k, o, v[k]);
    }
}

}  // namespace astro_dev
