// scan.h -- single-pass exclusive prefix sum with decoupled look-back, gfx950.
//
// Replaces the reference's cub::DeviceScan::ExclusiveSum call sites
// (src/sim/cell_store.cu:422, src/sim/lifecycle.cu:269) and the cell-range
// half of the hash build. An exclusive scan's output is uniquely defined, so
// any correct scan is bit-identical to CUB's.
//
// Design points:
//   * Tiles are handed out in increasing order by a CAS on a tile counter, so
//     a tile only ever waits on tiles whose blocks are already running: no
//     assumption about dispatch order, no deadlock.
//   * Tile status words carry a 32-bit generation ("epoch") read from the
//     device Counters block and advanced by the last tile. Nothing is ever
//     cleared between scans, so the scan needs no memset and can sit inside a
//     captured graph.
//   * The spin on a predecessor is bounded; on expiry it raises ERR_SCAN_SPIN
//     in Counters::error_flags (checked by the host) instead of hanging.
//   * Input sizes come from device memory, so the launch grid is sized for the
//     maximum and surplus blocks exit at once.
#pragma once

#include "astro_dev.h"

namespace astro_dev {

constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;
constexpr uint32_t SCAN_AGG = 1u, SCAN_INC = 2u;
constexpr uint32_t SCAN_SPIN_LIMIT = 1u << 22;

struct ScanCtl {
    unsigned long long* status;     // one word per tile, sized for the largest scan
    unsigned long long* tile_ctr;   // (epoch << 32) | next tile
    Counters* ctr;
};

__host__ __device__ __forceinline__ int scan_tiles_for(int n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

__device__ __forceinline__ unsigned long long scan_word(uint32_t epoch, uint32_t kind, uint32_t v) {
    return (static_cast<unsigned long long>(epoch) << 32) | (static_cast<unsigned long long>(kind) << 30) |
           static_cast<unsigned long long>(v & 0x3FFFFFFFu);
}

__device__ __forceinline__ int32_t wave_inclusive_sum(int32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Exclusive scan of load(i), i in [0, n), into store(i, prefix). The total is
// written to *total_out by the last tile (0 when n <= 0, by block 0).
// Every block of the launch must call this exactly once.
template <class Load, class Store>
__device__ void scan_exclusive(int n, Load load, Store store, const ScanCtl& sc, int32_t* total_out) {
    __shared__ uint32_t s_tile;
    __shared__ int32_t s_wave[BLOCK / 64];
    __shared__ int32_t s_prefix;
    if (n <= 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0 && total_out) *total_out = 0;
        return;
    }
    const int num_tiles = scan_tiles_for(n);
    if (static_cast<int>(blockIdx.x) >= num_tiles) return;
    const uint32_t epoch = sc.ctr->scan_epoch;
    if (threadIdx.x == 0) {
        unsigned long long old = __hip_atomic_load(sc.tile_ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t tile;
        for (;;) {
            unsigned long long want;
            if (static_cast<uint32_t>(old >> 32) != epoch) {
                tile = 0;
                want = (static_cast<unsigned long long>(epoch) << 32) | 1ull;
            } else {
                tile = static_cast<uint32_t>(old);
                want = old + 1ull;
            }
            const unsigned long long prev = atomicCAS(sc.tile_ctr, old, want);
            if (prev == old) break;
            old = prev;
        }
        s_tile = tile;
    }
    __syncthreads();
    const int tile = static_cast<int>(s_tile);
    const int base = tile * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int32_t v[SCAN_ITEMS];
    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int idx = base + k;
        v[k] = idx < n ? load(idx) : 0;
        sum += v[k];
    }
    const int32_t incl = wave_inclusive_sum(sum);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    int32_t wave_off = 0, block_total = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) {
        if (w < wave) wave_off += s_wave[w];
        block_total += s_wave[w];
    }
    if (threadIdx.x == 0) {
        int32_t excl = 0;
        if (tile == 0) {
            __hip_atomic_store(&sc.status[0], scan_word(epoch, SCAN_INC, static_cast<uint32_t>(block_total)),
                               __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&sc.status[tile], scan_word(epoch, SCAN_AGG, static_cast<uint32_t>(block_total)),
                               __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            int j = tile - 1;
            uint32_t spins = 0;
            for (;;) {
                const unsigned long long w =
                    __hip_atomic_load(&sc.status[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t kind = static_cast<uint32_t>(w >> 30) & 3u;
                if (static_cast<uint32_t>(w >> 32) == epoch && kind != 0u) {
                    excl += static_cast<int32_t>(w & 0x3FFFFFFFull);
                    if (kind == SCAN_INC) break;
                    --j;
                    continue;
                }
                if (++spins > SCAN_SPIN_LIMIT) {
                    atomicOr(&sc.ctr->error_flags, ERR_SCAN_SPIN);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __hip_atomic_store(&sc.status[tile], scan_word(epoch, SCAN_INC, static_cast<uint32_t>(excl + block_total)),
                               __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_prefix = excl;
        if (tile == num_tiles - 1) {
            if (total_out) *total_out = excl + block_total;
            sc.ctr->scan_epoch = epoch + 1u;
        }
    }
    __syncthreads();
    int32_t run = s_prefix + wave_off + incl - sum;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int idx = base + k;
        if (idx < n) store(idx, run);
        run += v[k];
    }
}

}  // namespace astro_dev
