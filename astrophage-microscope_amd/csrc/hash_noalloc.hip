// hash_noalloc.hip -- the stable spatial-hash build on caller-owned device
// arrays, for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The result is the reference's hash_build (src/sim/hash.cu:106-133) for the
// given cells: vals_sorted = the stable sort of slots by bucket key (free
// slots last, under the sentinel key nb), and each bucket's [start, end) in
// it, start == end == 0 when empty. Every intermediate lives in `scratch`, a
// file-scope device array each GPU holds for the life of the process; its
// device address is looked up once per device with hipGetSymbolAddress.
//
// Passes (all on the caller's stream, nothing waits for the host):
//   1. clear:   bucket counts and both per-tile total arrays to zero;
//   2. key:     bucket of each slot, arrival rank = atomicAdd on its count,
//               per-tile totals of the counts and of the free slots;
//   3. scan:    counts -> exclusive bucket starts, in place (scan.h);
//               free-slot flags -> their exclusive prefix (slot order);
//   4. scatter: each slot to its bucket start + arrival rank;
//   5. rank:    each slot re-ranked by slot inside its bucket (the stable
//               order; free slots take their prefix directly);
//   6. ranges:  start/end per bucket for the caller.
#include <mutex>

#include "scan.h"

namespace astro_dev {
namespace {

constexpr int32_t NA_CELLS = ASTRO_HASH_NOALLOC_MAX_CELLS;
constexpr int32_t NA_BUCKETS = ASTRO_HASH_NOALLOC_MAX_BUCKETS;

constexpr size_t round64(size_t w) { return (w + 63) / 64 * 64; }

// Word offsets into the scratch array.
constexpr size_t OFF_KEY = 0;                                        // [cells] bucket of each slot
constexpr size_t OFF_RANK = OFF_KEY + NA_CELLS;                      // [cells] arrival rank in its bucket
constexpr size_t OFF_ORDER = OFF_RANK + NA_CELLS;                    // [cells] slots in arrival order
constexpr size_t OFF_FREE = OFF_ORDER + NA_CELLS;                    // [cells] exclusive prefix of free slots
constexpr size_t OFF_START = OFF_FREE + NA_CELLS;                    // [buckets + 2] counts, then starts
constexpr size_t OFF_BTILES = OFF_START + round64(NA_BUCKETS + 2);   // per-scan-tile totals of the counts
constexpr size_t OFF_FTILES = OFF_BTILES + round64((NA_BUCKETS + 1 + SCAN_TILE - 1) / SCAN_TILE);
constexpr size_t OFF_ERR = OFF_FTILES + round64((NA_CELLS + SCAN_TILE - 1) / SCAN_TILE);
constexpr size_t NA_WORDS = OFF_ERR + 64;
static_assert(NA_WORDS * 4 == ASTRO_HASH_NOALLOC_SCRATCH_BYTES, "scratch size stated in astro_hot.h");

static __device__ __align__(256) uint32_t scratch[NA_WORDS];

struct NaView {
    uint32_t* key;
    uint32_t* rank;
    uint32_t* order;
    int32_t* free_prefix;
    int32_t* start;
    int32_t* btiles;
    int32_t* ftiles;
    uint32_t* err;
};

NaView view_of(uint32_t* base) {
    return NaView{base + OFF_KEY, base + OFF_RANK, base + OFF_ORDER,
                  reinterpret_cast<int32_t*>(base + OFF_FREE), reinterpret_cast<int32_t*>(base + OFF_START),
                  reinterpret_cast<int32_t*>(base + OFF_BTILES), reinterpret_cast<int32_t*>(base + OFF_FTILES),
                  base + OFF_ERR};
}

__global__ __launch_bounds__(BLOCK) void k_na_clear(NaView s, int32_t nbp1, int32_t btiles, int32_t ftiles) {
    const int stride = gridDim.x * BLOCK;
    for (int q = blockIdx.x * BLOCK + threadIdx.x; q < nbp1 + 1; q += stride) s.start[q] = 0;
    for (int q = blockIdx.x * BLOCK + threadIdx.x; q < btiles; q += stride) s.btiles[q] = 0;
    for (int q = blockIdx.x * BLOCK + threadIdx.x; q < ftiles; q += stride) s.ftiles[q] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *s.err = 0u;
}

// The trip count is uniform per wave, so every lane reaches the wave-wide
// tile counts together.
__global__ __launch_bounds__(BLOCK) void k_na_key(NaView s, const double* x, const double* y, const double* z,
                                                  const uint32_t* flags, int32_t n, HashGeo h) {
    for (int base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
        const int i = base + static_cast<int>(threadIdx.x);
        const bool valid = i < n;
        uint32_t key = 0u;
        bool occupied = false;
        if (valid) {
            occupied = (flags[i] & ASTRO_FLAG_OCCUPIED) != 0u;
            key = occupied ? hash_key(h, x[i], y[i], z[i]) : static_cast<uint32_t>(h.nb);
            s.key[i] = key;
            s.rank[i] = atomicAdd(reinterpret_cast<uint32_t*>(&s.start[key]), 1u);
        }
        wave_count_by_key(s.btiles, key >> SCAN_TILE_SHIFT, valid);
        wave_count_by_key(s.ftiles, static_cast<uint32_t>(i) >> SCAN_TILE_SHIFT, valid && !occupied);
    }
}

// Bucket counts -> exclusive starts in place; start[nb + 1] = n.
__global__ __launch_bounds__(BLOCK) void k_na_scan_buckets(NaView s, int32_t nbp1) {
    scan_tile(
        nbp1, blockIdx.x, [&](int i) { return load4_i32(s.start, i, nbp1); },
        [&](int i, int4 ex, int4) { store4_i32(s.start, i, nbp1, ex); }, s.btiles, s.start + nbp1, s.err);
}

// Free-slot flags (key == nb) -> their exclusive prefix in slot order.
__global__ __launch_bounds__(BLOCK) void k_na_scan_free(NaView s, int32_t n, uint32_t nb) {
    int32_t total = 0;
    scan_tile(
        n, blockIdx.x,
        [&](int i) {
            int4 v{0, 0, 0, 0};
            if (i < n) v.x = s.key[i] == nb;
            if (i + 1 < n) v.y = s.key[i + 1] == nb;
            if (i + 2 < n) v.z = s.key[i + 2] == nb;
            if (i + 3 < n) v.w = s.key[i + 3] == nb;
            return v;
        },
        [&](int i, int4 ex, int4) { store4_i32(s.free_prefix, i, n, ex); }, s.ftiles, &total, s.err);
}

__global__ __launch_bounds__(BLOCK) void k_na_scatter(NaView s, int32_t n) {
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        const int32_t dst = s.start[s.key[i]] + static_cast<int32_t>(s.rank[i]);
        if (dst >= 0 && dst < n) s.order[dst] = static_cast<uint32_t>(i);   // always, unless the scan failed
    }
}

// Stable order: ascending slot inside each bucket. Occupied buckets hold a
// few cells each (contact keeps cells apart), so a short count; free slots
// already have their rank from the scan.
__global__ __launch_bounds__(BLOCK) void k_na_rank(NaView s, int32_t n, uint32_t nb, uint32_t* vals_sorted) {
    for (int k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += gridDim.x * BLOCK) {
        const uint32_t slot = s.order[k];
        if (slot >= static_cast<uint32_t>(n)) continue;                     // only after a failed scan
        const uint32_t key = s.key[slot];
        const int32_t st = s.start[key];
        int32_t r;
        if (key == nb) {
            r = s.free_prefix[slot];
        } else {
            const int32_t lo = st < 0 ? 0 : st, hi = s.start[key + 1] < n ? s.start[key + 1] : n;
            r = 0;
            for (int32_t q = lo; q < hi; ++q) r += s.order[q] < slot ? 1 : 0;
        }
        if (st + r >= 0 && st + r < n) vals_sorted[st + r] = slot;   // always, unless the scan failed
    }
}

// A failed scan self-check leaves -1 in every bucket start, which no valid
// result holds.
__global__ __launch_bounds__(BLOCK) void k_na_ranges(NaView s, int32_t nb, int32_t* bstart, int32_t* bend) {
    const bool bad = *s.err != 0u;
    for (int b = blockIdx.x * BLOCK + threadIdx.x; b < nb; b += gridDim.x * BLOCK) {
        const int32_t st = s.start[b], en = s.start[b + 1];
        const bool any = en > st;
        bstart[b] = bad ? -1 : (any ? st : 0);
        bend[b] = any ? en : 0;
    }
}

int blocks_for(int64_t items) {
    const int64_t b = (items + BLOCK - 1) / BLOCK;
    return static_cast<int>(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

// The scratch array's device address on the current device, looked up once.
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

int32_t scratch_address(uint32_t** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return ASTRO_DEVICE_ERROR;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return ASTRO_DEVICE_ERROR;
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<uint32_t*>(g_scratch_ptr[dev]);
    return ASTRO_OK;
}

}  // namespace
}  // namespace astro_dev

extern "C" int32_t astro_hash_build_noalloc(const double* x, const double* y, const double* z, const uint32_t* flags,
                                            int32_t n, int32_t nx, int32_t ny, int32_t nz, double cell_size,
                                            uint32_t* vals_sorted, int32_t* bucket_start, int32_t* bucket_end,
                                            void* stream) {
    using namespace astro_dev;
    if (n < 0 || nx < 1 || ny < 1 || nz < 1 || !(cell_size > 0.0)) return ASTRO_INVALID_ARGUMENT;
    const int64_t nb64 = static_cast<int64_t>(nx) * ny * nz;
    if (n > NA_CELLS || nb64 > NA_BUCKETS) return ASTRO_CAPACITY_EXCEEDED;
    if ((n > 0 && (!x || !y || !z || !flags || !vals_sorted)) || !bucket_start || !bucket_end)
        return ASTRO_INVALID_ARGUMENT;
    uint32_t* base = nullptr;
    int32_t st = scratch_address(&base);
    if (st != ASTRO_OK) return st;
    const NaView s = view_of(base);
    const int32_t nb = static_cast<int32_t>(nb64);
    HashGeo h{};
    h.nx = nx; h.ny = ny; h.nz = nz; h.nb = nb;
    h.inv = static_cast<float>(1.0 / cell_size);
    h.ox = static_cast<float>(-0.5 * nx * cell_size);
    h.oy = static_cast<float>(-0.5 * ny * cell_size);
    h.oz = static_cast<float>(-0.5 * nz * cell_size);
    const hipStream_t sm = static_cast<hipStream_t>(stream);
    const int32_t btiles = scan_tiles_for(nb + 1), ftiles = scan_tiles_for(n);
    k_na_clear<<<blocks_for(nb + 2), BLOCK, 0, sm>>>(s, nb + 1, btiles, ftiles);
    if (n > 0) k_na_key<<<blocks_for(n), BLOCK, 0, sm>>>(s, x, y, z, flags, n, h);
    k_na_scan_buckets<<<btiles, BLOCK, 0, sm>>>(s, nb + 1);
    if (n > 0) {
        k_na_scan_free<<<ftiles, BLOCK, 0, sm>>>(s, n, static_cast<uint32_t>(nb));
        k_na_scatter<<<blocks_for(n), BLOCK, 0, sm>>>(s, n);
        k_na_rank<<<blocks_for(n), BLOCK, 0, sm>>>(s, n, static_cast<uint32_t>(nb), vals_sorted);
    }
    k_na_ranges<<<blocks_for(nb), BLOCK, 0, sm>>>(s, nb, bucket_start, bucket_end);
    return hipGetLastError() == hipSuccess ? ASTRO_OK : ASTRO_DEVICE_ERROR;
}
