// scan.h -- single-pass exclusive prefix sum with decoupled look-back, gfx950.
//
// Replaces the reference's cub::DeviceScan::ExclusiveSum call sites
// (src/sim/cell_store.cu:422, src/sim/lifecycle.cu:269) and the cell-range
// half of the hash build. An exclusive scan's output is uniquely defined, so
// any correct scan is bit-identical to CUB's.
//
// Design points:
//   * Tiles of 4096 items (256 threads x 16, read and written as 16-byte
//     vectors). Tile ids come from one atomicAdd on a monotonic counter minus
//     a device-held base, so a tile only ever waits on tiles whose blocks are
//     already running: no dispatch-order assumption, no deadlock, no CAS loop.
//   * Tile status words carry a 32-bit generation ("epoch"). The last tile
//     advances the epoch and the base, so nothing is cleared between scans:
//     no memset, and the scan can sit inside a captured graph.
//   * Look-back is done by one wave over a 512-tile window per probe.
//   * The spin on predecessors is bounded; on expiry it raises ERR_SCAN_SPIN
//     in Counters::error_flags (checked by the host) instead of hanging.
//   * Sizes come from device memory, so the grid is sized for the maximum and
//     surplus blocks exit at once.
#pragma once

#include "astro_dev.h"

namespace astro_dev {

constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;
constexpr uint32_t SCAN_AGG = 1u, SCAN_INC = 2u;
constexpr uint32_t SCAN_SPIN_LIMIT = 1u << 22;

struct ScanCtl {
    unsigned long long* status;     // one word per tile, sized for the largest scan
    unsigned long long* tile_ctr;   // monotonic tile counter
    Counters* ctr;                  // scan_epoch, scan_base, error_flags
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

__device__ __forceinline__ int32_t wave_sum(int32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Vector helpers for the common case of a plain int32 array.
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

// Exclusive scan of the items produced by load4(i) (items i..i+3, i a multiple
// of 4, zero past n) into store4(i, prefixes). The total goes to *total_out
// (from the last tile; 0 from block 0 when n <= 0). Every block of the launch
// must call this exactly once.
template <class Load4, class Store4>
__device__ void scan_exclusive(int n, Load4 load4, Store4 store4, const ScanCtl& sc, int32_t* total_out) {
    __shared__ int s_tile;
    __shared__ uint32_t s_epoch;
    __shared__ unsigned long long s_base;
    __shared__ int32_t s_wave[BLOCK / 64];
    __shared__ int32_t s_prefix;
    if (n <= 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0 && total_out) *total_out = 0;
        return;
    }
    const int num_tiles = scan_tiles_for(n);
    if (static_cast<int>(blockIdx.x) >= num_tiles) return;
    if (threadIdx.x == 0) {
        // Read the generation and base before taking a tile: the last tile
        // rewrites both once every tile has been taken.
        const uint32_t ep = __hip_atomic_load(&sc.ctr->scan_epoch, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long b =
            __hip_atomic_load(&sc.ctr->scan_base, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long t =
            __hip_atomic_fetch_add(sc.tile_ctr, 1ull, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_tile = static_cast<int>(t - b);
        s_epoch = ep;
        s_base = b;
    }
    __syncthreads();
    const int tile = s_tile;
    if (tile < 0 || tile >= num_tiles) {   // counters out of step: refuse, never index out of range
        if (threadIdx.x == 0) atomicOr(&sc.ctr->error_flags, ERR_SCAN_SPIN);
        return;
    }
    const uint32_t epoch = s_epoch;
    const unsigned long long base = s_base;
    const int first = tile * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int4 v[SCAN_ITEMS / 4];
    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS / 4; ++k) {
        v[k] = load4(first + 4 * k);
        sum += v[k].x + v[k].y + v[k].z + v[k].w;
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
    if (wave == 0) {
        int32_t excl = 0;
        if (tile == 0) {
            if (lane == 0)
                __hip_atomic_store(&sc.status[0], scan_word(epoch, SCAN_INC, static_cast<uint32_t>(block_total)),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0)
                __hip_atomic_store(&sc.status[tile], scan_word(epoch, SCAN_AGG, static_cast<uint32_t>(block_total)),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // Each probe covers 512 predecessors: lane l reads tiles
            // j - 8l - m, m = 0..7, nearest first, all loads in flight at once.
            int j = tile - 1;   // nearest predecessor not yet folded in
            uint32_t spins = 0;
            for (;;) {
                uint32_t kind[8], val[8];
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const int idx = j - 8 * lane - m;
                    kind[m] = SCAN_INC;   // before tile 0 reads as an inclusive 0
                    val[m] = 0u;
                    if (idx >= 0) {
                        const unsigned long long w =
                            __hip_atomic_load(&sc.status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        kind[m] = static_cast<uint32_t>(w >> 32) == epoch ? static_cast<uint32_t>(w >> 30) & 3u : 0u;
                        val[m] = static_cast<uint32_t>(w & 0x3FFFFFFFull);
                    }
                }
                int32_t lsum = 0;
                bool has_inc = false, ready = true;
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    if (has_inc || !ready) continue;
                    if (kind[m] == 0u) {
                        ready = false;
                        continue;
                    }
                    lsum += static_cast<int32_t>(val[m]);
                    if (kind[m] == SCAN_INC) has_inc = true;
                }
                const unsigned long long inc_mask = __ballot(has_inc);
                const unsigned long long not_ready = __ballot(!ready);
                const int stop = inc_mask ? __ffsll(static_cast<long long>(inc_mask)) - 1 : 64;
                const unsigned long long upto = stop >= 63 ? ~0ull : ((1ull << (stop + 1)) - 1ull);
                if ((not_ready & upto) == 0ull) {
                    excl += wave_sum(lane <= stop ? lsum : 0);
                    if (stop < 64) break;
                    j -= 512;
                    continue;
                }
                if (++spins > SCAN_SPIN_LIMIT) {
                    if (lane == 0) atomicOr(&sc.ctr->error_flags, ERR_SCAN_SPIN);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (lane == 0)
                __hip_atomic_store(&sc.status[tile], scan_word(epoch, SCAN_INC, static_cast<uint32_t>(excl + block_total)),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
            s_prefix = excl;
            if (tile == num_tiles - 1) {
                if (total_out) *total_out = excl + block_total;
                sc.ctr->scan_base = base + static_cast<unsigned long long>(num_tiles);
                sc.ctr->scan_epoch = epoch + 1u;
            }
        }
    }
    __syncthreads();
    int32_t run = s_prefix + wave_off + incl - sum;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS / 4; ++k) {
        int4 o;
        o.x = run; run += v[k].x;
        o.y = run; run += v[k].y;
        o.z = run; run += v[k].z;
        o.w = run; run += v[k].w;
        store4(first + 4 * k, o);
    }
}

}  // namespace astro_dev
