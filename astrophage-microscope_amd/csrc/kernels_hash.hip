// kernels_hash.hip -- spatial-hash build and the two neighbour walks.
//
// The reference builds the hash with a CUB stable radix sort of (bucket, slot)
// pairs plus an adjacent-key range pass (src/sim/hash.cu:106-133). Here the
// key *is* the sort digit, so the build is a counting sort over the bucket
// index (a one-digit LSD radix sort):
//   1. key:     bucket of each slot; arrival rank = atomicAdd on its count.
//   2. scan:    dense exclusive prefix of the counts (one block per tile,
//               from per-tile totals the key pass accumulated); the counts
//               are zeroed in the same pass for the next build.
//   3. scatter: records to bucket base + arrival rank (order inside a bucket
//               is still arrival order, i.e. scheduling dependent).
//   4. rank:    each record re-ranks itself by slot inside its bucket, which
//               makes the result exactly the stable sort: ascending slot
//               within ascending bucket, as the reference requires.
// Buckets hold a handful of cells (contact keeps them apart), so step 4 is a
// short local loop. The sorted records carry positions, so the walks read
// neighbours contiguously in bucket order instead of gathering by slot.
// Cells live physically in the order of the last re-sort, so every pass over
// them here is a near-sequential stream; the sort key's tie-break is the
// reference slot (lslot), which is what makes the order the reference's.
#include <algorithm>

#include "kernels.h"

namespace astro_dev {

namespace {

template <class P>
struct KHashKey {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> c;
        Counters* ctr;
        HashGeo h;
        uint32_t* keys;
        uint32_t* rin;
        uint32_t* bcount;
        int32_t* bsum;
        int32_t* dflag;
        int32_t* drop;
        int32_t* dsum;
        int32_t* xsum;
        int cap;
        int count_tick;
        int resort_period;
        Grid T;            // with tcount: the temperature grid whose voxels are counted
        uint32_t* tcount;  // non-null: also count the occupied entries per T voxel
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [c, ctr, h, keys, rin, bcount, bsum, dflag, drop, dsum, xsum, cap, count_tick, resort_period, T,
                     tcount] = a;

    const int np = ctr->nphys;
    const int gid = vg.block * BLOCK + threadIdx.x, stride = vg.grid * BLOCK;
    if (count_tick) {
        if (gid == 0) {
            const int n = ctr->count;
            ctr->cell_steps += n;
            const uint32_t t = ctr->tick_ctr;
            ctr->resort = (t % static_cast<uint32_t>(resort_period) == 0u || np > cap) ? 1 : 0;
            ctr->tick_ctr = t + 1u;
            // the tick in progress, for counter-based streams
            const uint64_t k = ctr->tick_count;
            ctr->rng_tick = k;
            ctr->tick_count = k + 1u;
        }
        // The lifecycle's per-slot flags and their scan-tile totals start
        // each tick cleared.
        for (int q = gid; q < cap; q += stride) {
            dflag[q] = 0;
            drop[q] = 0;
        }
        for (int q = gid; q < scan_tiles_for(cap); q += stride) {
            dsum[q] = 0;
            xsum[q] = 0;
        }
    }
    xcd_tiles(vg, np, [&](const int p, const bool valid) {
        // Slot, flags and position are loaded together (all in bounds for p
        // below nphys), so a single memory latency precedes the key.
        uint32_t s = NO_SLOT, f = 0u;
        P x{}, y{}, z{};
        if (valid) {
            s = c.lslot[p];
            f = c.flags[p];
            x = c.x[p];
            y = c.y[p];
            z = c.z[p];
        }
        const bool live = s != NO_SLOT;
        uint32_t key = 0u;
        if (live) {
            key = (f & ASTRO_FLAG_OCCUPIED)
                ? hash_key(h, static_cast<double>(x), static_cast<double>(y), static_cast<double>(z))
                : static_cast<uint32_t>(h.nb);
            keys[p] = key;
            rin[p] = atomicAdd(&bcount[key], 1u);
            // the thermostat's voxel count (KThermalCount), from the same
            // positions: nothing moves a cell before the thermal step
            if (tcount && (f & ASTRO_FLAG_OCCUPIED))
                atomicAdd(&tcount[grid_nearest(T, static_cast<float>(x), static_cast<float>(y), static_cast<float>(z))],
                          1u);
        }
        wave_count_by_key(bsum, key >> SCAN_TILE_SHIFT, live);
    });
    }
};

struct KBucketScan {
    struct Args {
        Counters* ctr;
        uint32_t* bcount;
        const int32_t* bsum;
        int32_t* bstart;
        int nbp1;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [ctr, bcount, bsum, bstart, nbp1] = a;

    int32_t* counts = reinterpret_cast<int32_t*>(bcount);
    scan_tile(
        nbp1, vg.block,
        [&](int i) {
            const int4 v = load4_i32(counts, i, nbp1);
            if (v.x | v.y | v.z | v.w) store4_i32(counts, i, nbp1, int4{0, 0, 0, 0});
            return v;
        },
        [&](int i, int4 ex, int4) { store4_i32(bstart, i, nbp1, ex); }, bsum, bstart + nbp1, &ctr->error_flags);
    }
};

template <class P>
struct KHashScatter {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> c;
        const Counters* ctr;
        const uint32_t* keys;
        const uint32_t* rin;
        const int32_t* bstart;
        Rec<P>* trec;
        uint32_t* tphys;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [c, ctr, keys, rin, bstart, trec, tphys] = a;

    const int np = ctr->nphys;
    const auto body = [&](const int p) {
    const uint32_t s = c.lslot[p];
    if (s == NO_SLOT) return;
    const uint32_t f = c.flags[p];
    Rec<P> r;
    r.x = c.x[p];
    r.y = c.y[p];
    r.z = c.z[p];
    r.meta = s | ((f & ASTRO_FLAG_OCCUPIED) ? 0u : REC_UNOCC) | ((f & ASTRO_FLAG_ALIVE) ? REC_ALIVE : 0u);
    if constexpr (sizeof(Rec<P>) == 32) r._pad = 0u;
    const int32_t dst = bstart[keys[p]] + static_cast<int32_t>(rin[p]);
    trec[dst] = r;
    tphys[dst] = static_cast<uint32_t>(p);
    };
    xcd_tiles(vg, np, [&](const int p, const bool valid) { if (valid) body(p); });
    }
};

// Stable order: ascending reference slot inside each bucket.
template <class P>
struct KHashRank {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        const Counters* ctr;
        HashGeo h;
        const Rec<P>* trec;
        const uint32_t* tphys;
        const int32_t* bstart;
        Rec<P>* rec;
        uint32_t* vals_sorted;
        uint32_t* sphys;
        int32_t* spos;
        int32_t* bsum;
        int bsum_tiles;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [ctr, h, trec, tphys, bstart, rec, vals_sorted, sphys, spos, bsum, bsum_tiles] = a;

    const int n = ctr->count;
    // The bucket scan has consumed the tile totals; clear them for the next
    // build's key pass.
    for (int q = vg.block * BLOCK + threadIdx.x; q < bsum_tiles; q += vg.grid * BLOCK) bsum[q] = 0;
    const auto body = [&](const int k) {
    const Rec<P> r = trec[k];
    const uint32_t slot = r.meta & REC_SLOT_MASK;
    const uint32_t key = (r.meta & REC_UNOCC)
        ? static_cast<uint32_t>(h.nb)
        : hash_key(h, static_cast<double>(r.x), static_cast<double>(r.y), static_cast<double>(r.z));
    const int32_t s = bstart[key], e = bstart[key + 1];
    int32_t rank = 0;
    for (int32_t q = s; q < e; ++q) rank += ((trec[q].meta & REC_SLOT_MASK) < slot) ? 1 : 0;
    const int32_t dst = s + rank;
    const uint32_t p = tphys[k];
    rec[dst] = r;
    vals_sorted[dst] = slot;
    sphys[dst] = p;
    spos[p] = dst;
    };
    for (int k = vg.block * BLOCK + threadIdx.x; k < n; k += vg.grid * BLOCK) body(k);
    }
};

// Re-sort: move every live entry to its sorted position (drops tombstones),
// then copy back. Runs only on ticks the key kernel marked.
template <class P>
struct KResortGather {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> live;
        Cells<P> alt;
        const Counters* ctr;
        const uint32_t* sphys;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [live, alt, ctr, sphys] = a;

    if (!ctr->resort) return;
    const int n = ctr->count;
    const auto body = [&](const int k) {
    copy_cell(live, static_cast<int>(sphys[k]), alt, k);
    };
    for (int k = vg.block * BLOCK + threadIdx.x; k < n; k += vg.grid * BLOCK) body(k);
    }
};

template <class P>
struct KResortCopyback {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> live;
        Cells<P> alt;
        Counters* ctr;
        uint32_t* sphys;
        int32_t* spos;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [live, alt, ctr, sphys, spos] = a;

    if (!ctr->resort) return;
    const int n = ctr->count;
    // No block of this pass reads nphys, so one thread can already publish
    // the packed extent for the kernels that follow.
    if (vg.block == 0 && threadIdx.x == 0) ctr->nphys = n;
    const auto body = [&](const int k) {
    copy_cell(alt, k, live, k);
    sphys[k] = static_cast<uint32_t>(k);
    spos[k] = k;
    };
    for (int k = vg.block * BLOCK + threadIdx.x; k < n; k += vg.grid * BLOCK) body(k);
    }
};

// Record range [s, e) of row `row` (0..8) of the 27 buckets around bucket
// (gx, gy, gz), the rows in the reference's walk order (z, then y, then x,
// then sorted position, src/sim/hash.cuh:88-111). The three x-adjacent
// buckets of one (y, z) row are contiguous in the sorted array, so a row is
// the single range [bstart[x_lo], bstart[x_hi + 1]); a row off the grid is
// empty.
__device__ __forceinline__ void row_range(const HashGeo& h, const int32_t* bstart, int gx, int gy, int gz, int row,
                                          int& s, int& e) {
    const int xlo = gx > 0 ? gx - 1 : 0;
    const int xhi = gx < h.nx - 1 ? gx + 1 : h.nx - 1;
    const int cz = gz + row / 3 - 1, cy = gy + row % 3 - 1;
    const bool ok = cz >= 0 && cz < h.nz && cy >= 0 && cy < h.ny;
    const int base = ok ? (cz * h.ny + cy) * h.nx : 0;
    s = bstart[base + xlo];
    e = ok ? bstart[base + xhi + 1] : s;
}

// Position difference, as the float the contact screen uses: exact to one
// float rounding of the double difference the reference computes.
__device__ __forceinline__ float screen_diff(float a, float b) { return a - b; }
__device__ __forceinline__ float screen_diff(double a, double b) { return static_cast<float>(a - b); }

// (2a)^2 (1 + 2^-12): the shadow screen's bound on a neighbour's squared
// distance from the light line.
constexpr float SHADOW_SCREEN = static_cast<float>(4.0 * ASTRO_CELL_RADIUS * ASTRO_CELL_RADIUS * (1.0 + 0x1p-12));

// The shadow walk's float screen: true for a live neighbour o that may shade
// the cell r. A neighbour downstream of the cell (along <= 0) or at least 2a
// off its light line has an overlap of exactly 0 and multiplies the open
// fraction by exactly 1. The float along and perp^2 differ from the double
// ones by under 2^-21 (|dx|+|dy|+|dz|) and 2^-18 |d|^2, so these margins never
// drop a shading neighbour.
template <class P>
__device__ __forceinline__ bool shadow_hit(const Rec<P>& r, const Rec<P>& o, float lxf, float lyf, float lzf) {
    const float sx = screen_diff(r.x, o.x), sy = screen_diff(r.y, o.y), sz = screen_diff(r.z, o.z);
    const float sa = sx * lxf + sy * lyf + sz * lzf;
    const float s2 = sx * sx + sy * sy + sz * sz;
    const bool upstream = !(sa < -0x1p-16f * (fabsf(sx) + fabsf(sy) + fabsf(sz)));
    const bool near_line = !(s2 - sa * sa > SHADOW_SCREEN + 0x1p-14f * s2);
    return ((o.meta & REC_ALIVE) != 0) & upstream & near_line;
}

// Records copied as raw 16-byte words (a local array of Rec<P> ended up in
// per-lane scratch memory in the batched walks; vectors stay in registers).
template <class P>
struct RecWords {
    uint4 w[sizeof(Rec<P>) / 16];
};
template <class P>
__device__ __forceinline__ RecWords<P> load_rec(const Rec<P>* p) {
    RecWords<P> r;
#pragma unroll
    for (int q = 0; q < static_cast<int>(sizeof(Rec<P>) / 16); ++q) r.w[q] = reinterpret_cast<const uint4*>(p)[q];
    return r;
}
template <class P>
__device__ __forceinline__ void store_rec(Rec<P>* p, const RecWords<P>& r) {
#pragma unroll
    for (int q = 0; q < static_cast<int>(sizeof(Rec<P>) / 16); ++q) reinterpret_cast<uint4*>(p)[q] = r.w[q];
}

// Possible shading neighbours a thread queues before the exact pass.
constexpr int SHADOW_QUEUE = 8;

// Possible contacts a thread queues before the exact pass (the queue is
// drained early, in order, if it fills).
constexpr int CONTACT_QUEUE = 4;


// Records staged per wave and z offset (6 KB of LDS per wave).
template <class P>
constexpr int walk_seg() { return (6 * 1024) / static_cast<int>(sizeof(Rec<P>)); }

// Steps of its walk a lane screens per round: one bit of its survivor mask
// each.
constexpr int WALK_ROUND = 32;

// The neighbour walk of a whole wave, in masked rounds. Each active lane
// visits the records of the 27 buckets around its own (gx, gy, gz) in the
// reference's walk order (z, y, x, then sorted position,
// src/sim/hash.cuh:88-111), skipping its own record (sorted position k);
// every lane of the wave must call it.
//   * Per z offset the wave stages the records of its three rows into its
//     LDS slice `seg`: the 64 consecutive records of a wave have keys in
//     [kmin, kmax], and keys are linear in (z, y, x), so row r of every
//     lane's walk lies in one contiguous record range. The three ranges sit
//     end to end ("positions" 0 .. total) and are copied in chunks of
//     walk_seg records with coalesced loads. Each lane then walks its own
//     three rows back to back in one loop; the wave runs as many iterations
//     as its longest three-row walk, and the loop body has no branch.
//   * hit(o) screens neighbour o. A lane marks the steps that pass in a
//     32-bit mask, and after each round of at most WALK_ROUND steps queues
//     those records' slice indices in walk order (QUEUE entries per lane,
//     entry j at queue[j * BLOCK]).
//   * term(o) takes the queued records in walk order, reading them from the
//     slice: when the queue fills, and at the end of each chunk, before the
//     next chunk reuses the slice. So a lane sees its survivors in the order
//     of the reference's walk.
// Returns true if the lane queued any record.
template <class P, int QUEUE, class Hit, class Term>
__device__ __forceinline__ bool masked_walk(const HashGeo& h, const Rec<P>* rec, const int32_t* bstart, Rec<P>* seg,
                                            int32_t* queue, bool active, int k, int gx, int gy, int gz, Hit&& hit,
                                            Term&& term) {
    constexpr int SEG = walk_seg<P>();
    const int lane = static_cast<int>(threadIdx.x) & 63;
    int queued = 0;
    bool any = false;
    const auto drain = [&]() {
        any = true;
        for (int j = 0; j < queued; ++j) term(seg[queue[j * BLOCK]]);
        queued = 0;
    };
    // Row (dz, dy) of any lane lies in buckets [kmin + o - 1, kmax + o + 1],
    // o = dz*nx*ny + dy*nx. Lane b < 18 fetches the bucket start of bound
    // b (the start of row b/2's range, or the end), and the wave reads
    // them back by lane.
    const int xlo = gx > 0 ? gx - 1 : 0;
    const int xhi = gx < h.nx - 1 ? gx + 1 : h.nx - 1;
    const int key = (gz * h.ny + gy) * h.nx + gx;
    // The wave's records are consecutive in bucket order, unoccupied ones
    // (the sentinel bucket) last: its active lanes hold keys that do not
    // decrease with the lane, so the first and last of them bound the rest.
    const unsigned long long act = __ballot(active);
    const int kmin = act ? __builtin_amdgcn_readlane(key, __ffsll(static_cast<long long>(act)) - 1) : 0x7FFFFFFF;
    const int kmax = act ? __builtin_amdgcn_readlane(key, 63 - __clzll(static_cast<long long>(act))) : -1;
    int bound = 0;
    if (kmax >= 0 && lane < 18) {
        const int row = lane >> 1;
        const int64_t o = static_cast<int64_t>(row / 3 - 1) * h.nx * h.ny + static_cast<int64_t>(row % 3 - 1) * h.nx;
        int64_t b = (lane & 1) ? static_cast<int64_t>(kmax) + o + 2 : static_cast<int64_t>(kmin) + o - 1;
        b = b < 0 ? 0 : (b > h.nb ? h.nb : b);
        bound = bstart[b];
    }
#pragma unroll 1
    for (int d = 0; d < 3; ++d) {
        // The wave's three ranges [lo[j], lo[j] + len) at positions off[j].
        int lo[3], off[3], total = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int ra = __builtin_amdgcn_readlane(bound, 6 * d + 2 * j);
            const int rb = __builtin_amdgcn_readlane(bound, 6 * d + 2 * j + 1);
            lo[j] = ra;
            off[j] = total;
            total += rb > ra ? rb - ra : 0;
        }
        // This lane's rows (dz = d - 1, dy = j - 1), as row_ranges computes
        // them: nr[j] records from position p[j]. Each lies inside the
        // wave's range of its row, so positions rise along the lane's walk.
        // The six loads are unconditional (a row off the grid reads row 0,
        // which is in bounds, and counts no records), so they issue together
        // with no branch.
        int p[3], nr[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int cz = gz + d - 1, cy = gy + j - 1;
            const bool ok = active && cz >= 0 && cz < h.nz && cy >= 0 && cy < h.ny;
            const int row = ok ? (cz * h.ny + cy) * h.nx : 0;
            const int rs = bstart[row + xlo];
            const int re = bstart[row + xhi + 1];
            nr[j] = ok && re > rs ? re - rs : 0;
            p[j] = off[j] + (rs - lo[j]);
        }
        const int n01 = nr[0] + nr[1];
        // The lane's own record, in the middle row of the middle z offset.
        const int selfpos = (d == 1 && active) ? off[1] + (k - lo[1]) : -1;
        // Position of step m of the lane's walk.
        const auto pos_of = [&](int m) {
            return m < nr[0] ? p[0] + m : (m < n01 ? p[1] + (m - nr[0]) : p[2] + (m - n01));
        };
        int m = 0;
        for (int c0 = 0; c0 < total; c0 += SEG) {
            const int cend = total - c0 < SEG ? total : c0 + SEG;
            // Up to four records per lane are loaded before any is stored,
            // so the copy costs one memory latency, not four. Loads and
            // stores are unconditional: a position past the chunk is clamped
            // to its last one, whose record is read and written again
            // unchanged.
            for (int v = c0 + lane; v < cend; v += 4 * 64) {
                RecWords<P> t[4];
                int w[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    w[u] = min(v + 64 * u, cend - 1);
                    const int shift =
                        w[u] < off[1] ? lo[0] - off[0] : (w[u] < off[2] ? lo[1] - off[1] : lo[2] - off[2]);
                    t[u] = load_rec(rec + (w[u] + shift));
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) store_rec(seg + (w[u] - c0), t[u]);
            }
            wave_sync();
            // The lane's steps whose positions lie below cend.
            int mc = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int below = cend - p[j];
                mc += below < 0 ? 0 : (below < nr[j] ? below : nr[j]);
            }
            while (__ballot(m < mc) != 0) {
                const int sb = m;
                const int mend = mc - sb < WALK_ROUND ? mc : sb + WALK_ROUND;
                uint32_t mask = 0;
                int pos = pos_of(m);
                // Position after step q at position at: the next in its row,
                // or the start of the next row.
                const auto next_pos = [&](int at, int q) {
                    int nx_ = at + 1;
                    if (q + 1 == nr[0]) nx_ = p[1];
                    if (q + 1 == n01) nx_ = p[2];
                    return nx_;
                };
                const auto hit_at = [&](int at) {
                    return static_cast<uint32_t>(hit(seg[at - c0]) && at != selfpos);
                };
                // Two steps per iteration: both records are read before
                // either is screened, so two LDS reads share one wait.
                for (; m + 1 < mend; m += 2) {
                    const int a0 = pos, a1 = next_pos(a0, m);
                    pos = next_pos(a1, m + 1);
                    const uint32_t h0 = hit_at(a0), h1 = hit_at(a1);
                    mask |= (h0 | (h1 << 1)) << (m - sb);
                }
                if (m < mend) {
                    mask |= hit_at(pos) << (m - sb);
                    ++m;
                }
                // Queue the round's survivors by slice index, in order.
                while (mask != 0u) {
                    const int at = pos_of(sb + __builtin_ctz(mask));
                    mask &= mask - 1u;
                    if (queued == QUEUE) drain();
                    queue[queued * BLOCK] = at - c0;
                    ++queued;
                }
            }
            // The queue refers to this chunk's slice: drain it before the
            // next chunk reuses the slice.
            if (queued > 0) drain();
            wave_sync();
        }
    }
    return any;
}

// shadow_fraction's factor (src/sim/emission.cuh:30-37) for neighbour o of
// the cell at (px, py, pz): 1 - the overlap of the two discs seen along the
// light, 1.0 for a neighbour downstream.
template <class P>
__device__ __forceinline__ double shadow_factor(double px, double py, double pz, const Rec<P>& o, double ldx,
                                                double ldy, double ldz) {
    const double dx = px - static_cast<double>(o.x);
    const double dy = py - static_cast<double>(o.y);
    const double dz = pz - static_cast<double>(o.z);
    const double along = dx * ldx + dy * ldy + dz * ldz;
    double frac = 0.0;
    if (along > 0.0) {
        const double perp2 = (dx * dx + dy * dy + dz * dz) - along * along;
        const double perp = perp2 > 0.0 ? sqrt(perp2) : 0.0;
        frac = disc_overlap_fraction(perp, ASTRO_CELL_RADIUS);
    }
    return 1.0 - frac;
}

// The contact half of forces_kernel (src/sim/integrator.cu:60-67), one lane
// per sorted record, adding to the body force already in fs[k]. Neighbour
// positions come from the hash's records, fixed since the hash build; the
// walk is masked_walk's.
//   * Each neighbour is screened in float: the float r^2 of a pair (the
//     differences rounded to float, squared and summed with fused
//     multiply-adds) is within a few float roundings (relative 2^-21) of the
//     double r^2, so a 2^-10 margin keeps every pair the reference counts as
//     touching.
//   * The survivors go through contact_force in walk order, so F sums the
//     same terms in the same order as the reference: in double in the
//     contract layout, bit for bit, and in float in the FP32 layout, whose
//     force buffer and integrator are float (each term then differs from the
//     double one by a few float roundings, at most about k * 2^-21 * 2a for
//     pairs at the contact threshold). Pairs the exact test finds apart add
//     a zero vector; F's components are sums of nonzero terms onto a nonzero
//     or +0.0 start and so never -0.0, which makes those adds exact no-ops.
template <class P>
struct KContactWalk {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        const Counters* ctr;
        HashGeo h;
        const Rec<P>* rec;
        const int32_t* bstart;
        Force<P>* fs;
        double k_contact;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [ctr, h, rec, bstart, fs, k_contact] = a;
        __shared__ Rec<P> s_seg[BLOCK / 64][walk_seg<P>()];
        __shared__ int32_t s_queue[CONTACT_QUEUE * BLOCK];
        Rec<P>* seg = s_seg[threadIdx.x >> 6];
        int32_t* queue = s_queue + threadIdx.x;
        const int n = ctr->count;
        const double two_a = ASTRO_CELL_DIAMETER;
        const float screen = static_cast<float>(two_a * two_a) * (1.0f + 1.0f / 1024.0f);
        xcd_tiles(vg, n, [&](const int k, const bool valid) {
            Rec<P> r{};
            bool occ = false;
            int gx = 0, gy = 0, gz = 0;
            if (valid) {
                r = rec[k];
                occ = !(r.meta & REC_UNOCC);
                if (occ)
                    hash_coords(h, static_cast<double>(r.x), static_cast<double>(r.y), static_cast<double>(r.z), gx,
                                gy, gz);
            }
            // The body force the sample stored, loaded now so that its latency
            // overlaps the walk; it is written back only if a contact was
            // queued.
            Force<P> F{};
            if (occ) F = fs[k];
            // contact_force (src/sim/integrator.cuh:210-219) in the layout's width
            const P two_a_p = static_cast<P>(two_a), k_p = static_cast<P>(k_contact);
            const bool touched = masked_walk<P, CONTACT_QUEUE>(
                h, rec, bstart, seg, queue, occ, k, gx, gy, gz,
                [&](const Rec<P>& o) {
                    const float dx = screen_diff(r.x, o.x), dy = screen_diff(r.y, o.y), dz = screen_diff(r.z, o.z);
                    // Fused multiply-adds round fewer times than separate
                    // products and sums, well inside the margin.
                    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                    return r2 < screen;
                },
                [&](const Rec<P>& o) {
                    const P dx = r.x - o.x;
                    const P dy = r.y - o.y;
                    const P dz = r.z - o.z;
                    const P r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 >= two_a_p * two_a_p || r2 <= P(0)) return;
                    const P rr2 = sqrt(r2);
                    const P overlap = two_a_p - rr2;
                    const P s = k_p * overlap / rr2;
                    F.x += dx * s;
                    F.y += dy * s;
                    F.z += dz * s;
                });
            if (touched) fs[k] = F;
        });
    }
};

// feed_kernel's irradiance half (src/sim/emission.cu:81-105): far-field sample
// times the exact near-field shadow product over upstream live neighbours,
// walked as the contact walk walks (masked_walk).
//   * The screen is shadow_hit's, in float.
//   * The survivors' factors are evaluated in double and multiplied into
//     `open` in walk order, as the reference multiplies them.
template <class P>
struct KFeedWalk {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        const Counters* ctr;
        HashGeo h;
        const Rec<P>* rec;
        const int32_t* bstart;
        Grid irr;
        double ldx;
        double ldy;
        double ldz;
        float ambient;
        int occl;
        double* dirs;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [ctr, h, rec, bstart, irr, ldx, ldy, ldz, ambient, occl, dirs] = a;
        const float lxf = static_cast<float>(ldx), lyf = static_cast<float>(ldy), lzf = static_cast<float>(ldz);
        __shared__ Rec<P> s_seg[BLOCK / 64][walk_seg<P>()];
        __shared__ int32_t s_queue[SHADOW_QUEUE * BLOCK];
        Rec<P>* seg = s_seg[threadIdx.x >> 6];
        int32_t* queue = s_queue + threadIdx.x;
        const int n = ctr->count;
        xcd_tiles(vg, n, [&](const int k, const bool valid) {
            Rec<P> r{};
            bool lit = false;
            if (valid) {
                r = rec[k];
                lit = !(r.meta & REC_UNOCC) && (r.meta & REC_ALIVE);
            }
            const double px = static_cast<double>(r.x), py = static_cast<double>(r.y), pz = static_cast<double>(r.z);
            double directional = 0.0;
            if (lit) {
                directional = static_cast<double>(grid_sample(irr, static_cast<float>(r.x), static_cast<float>(r.y),
                                                              static_cast<float>(r.z))) -
                              static_cast<double>(ambient);
                if (directional < 0.0) directional = 0.0;
            }
            if (occl) {
                double open = 1.0;
                int gx = 0, gy = 0, gz = 0;
                if (lit) hash_coords(h, px, py, pz, gx, gy, gz);
                masked_walk<P, SHADOW_QUEUE>(
                    h, rec, bstart, seg, queue, lit, k, gx, gy, gz,
                    [&](const Rec<P>& o) { return shadow_hit(r, o, lxf, lyf, lzf); },
                    [&](const Rec<P>& o) { open *= shadow_factor(px, py, pz, o, ldx, ldy, ldz); });
                directional *= open;
            }
            if (lit) dirs[k] = directional;
        });
    }
};

// The feed walk for crowded worlds (several cells per hash bucket, so each
// cell screens on the order of a hundred neighbours): one 16-lane group per
// cell, lane r < 9 walking row r of the 27-bucket neighbourhood (the rows in
// the reference's walk order, each in sorted position order). A lane screens
// its row as KFeedWalk does and evaluates its survivors' factors in double,
// keeping them in order in LDS; the group's first lane then multiplies all of
// them into `open` row after row, in the same order and with the same factors
// as the one-lane walk, so the product is the same bit for bit. A row with
// more survivors than the LDS keeps is walked again, in order, by the first
// lane itself. The neighbour records are read from the hash array directly,
// one dependent load after another, so the world's caller takes this walk
// only where all the launch's records fit one XCD's 4 MiB L2.
constexpr int FEED_GROUP = 16;
constexpr int FEED_ROW_CAP = 24;

template <class P>
struct KFeedWalkRows {
    static constexpr bool stride = true;   // any block count covers the work
    using Args = typename KFeedWalk<P>::Args;
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [ctr, h, rec, bstart, irr, ldx, ldy, ldz, ambient, occl, dirs] = a;
        constexpr int G = FEED_GROUP, CELLS = BLOCK / FEED_GROUP;
        __shared__ double s_f[CELLS][9][FEED_ROW_CAP];
        __shared__ int s_n[CELLS][9];
        const float lxf = static_cast<float>(ldx), lyf = static_cast<float>(ldy), lzf = static_cast<float>(ldz);
        const int j = static_cast<int>(threadIdx.x) % G, grp = static_cast<int>(threadIdx.x) / G;
        const int n = ctr->count;
        for (int k0 = vg.block * CELLS; k0 < n; k0 += vg.grid * CELLS) {   // uniform over the block
            const int k = k0 + grp;
            Rec<P> r{};
            bool lit = false;
            if (k < n) {
                r = rec[k];
                lit = !(r.meta & REC_UNOCC) && (r.meta & REC_ALIVE);
            }
            const uint32_t self = r.meta & REC_SLOT_MASK;
            const double px = static_cast<double>(r.x), py = static_cast<double>(r.y), pz = static_cast<double>(r.z);
            int gx = 0, gy = 0, gz = 0;
            if (lit) hash_coords(h, px, py, pz, gx, gy, gz);
            const auto screen = [&](const Rec<P>& o) {
                return (o.meta & REC_SLOT_MASK) != self && shadow_hit(r, o, lxf, lyf, lzf);
            };
            if (occl && lit && j < 9) {
                int s, e, cnt = 0;
                row_range(h, bstart, gx, gy, gz, j, s, e);
                for (int q = s; q < e; ++q) {
                    const Rec<P> o = rec[q];
                    if (!screen(o)) continue;
                    if (cnt < FEED_ROW_CAP) s_f[grp][j][cnt] = shadow_factor(px, py, pz, o, ldx, ldy, ldz);
                    ++cnt;
                }
                s_n[grp][j] = cnt;
            }
            __syncthreads();
            if (j == 0 && lit) {
                double directional = static_cast<double>(grid_sample(irr, static_cast<float>(r.x),
                                                                     static_cast<float>(r.y), static_cast<float>(r.z))) -
                                     static_cast<double>(ambient);
                if (directional < 0.0) directional = 0.0;
                if (occl) {
                    double open = 1.0;
                    for (int row = 0; row < 9; ++row) {
                        const int cnt = s_n[grp][row];
                        if (cnt <= FEED_ROW_CAP) {
                            for (int i = 0; i < cnt; ++i) open *= s_f[grp][row][i];
                        } else {
                            int s, e;
                            row_range(h, bstart, gx, gy, gz, row, s, e);
                            for (int q = s; q < e; ++q) {
                                const Rec<P> o = rec[q];
                                if (screen(o)) open *= shadow_factor(px, py, pz, o, ldx, ldy, ldz);
                            }
                        }
                    }
                    directional *= open;
                }
                dirs[k] = directional;
            }
            __syncthreads();   // the slots are reused by the next cells
        }
    }
};

}  // namespace

template <class P>
void launch_hash_build(const Launch& L, Views<P> ws, bool count_tick, bool thermal_count) {
    Batch<KHashKey<P>> key;
    Batch<KBucketScan> scan;
    Batch<KHashScatter<P>> scatter;
    Batch<KHashRank<P>> rank;
    for (const WorldDev<P>& w : ws) {
        const int gp = cell_grid(w.pcap);
        const int nbp1 = w.hg.nb + 1;
        key.add({w.c, w.ctr, w.hg, w.keys, w.rin, w.bcount, w.bsum, w.dflag, w.drop, w.dsum, w.xsum, w.capacity,
                 count_tick ? 1 : 0, w.resort_period, w.g[ASTRO_FIELD_TEMPERATURE], thermal_count ? w.tcount : nullptr},
                gp);
        scan.add({w.ctr, w.bcount, w.bsum, w.bstart, nbp1}, scan_tiles_for(nbp1));
        scatter.add({w.c, w.ctr, w.keys, w.rin, w.bstart, w.trec, w.tphys}, gp);
        rank.add({w.ctr, w.hg, w.trec, w.tphys, w.bstart, w.rec, w.vals_sorted, w.sphys, w.spos, w.bsum,
                  scan_tiles_for(nbp1)},
                 stride_grid(w.capacity));
    }
    launch_k<KHashKey<P>>(L, "hash_key", key.args, key.blocks);
    launch_k<KBucketScan>(L, "hash_bucket_scan", scan.args, scan.blocks);
    launch_k<KHashScatter<P>>(L, "hash_scatter", scatter.args, scatter.blocks);
    launch_k<KHashRank<P>>(L, "hash_rank", rank.args, rank.blocks);
}

template <class P>
void launch_resort(const Launch& L, Views<P> ws) {
    Batch<KResortGather<P>> gather;
    Batch<KResortCopyback<P>> copyback;
    for (const WorldDev<P>& w : ws) {
        gather.add({w.c, w.alt, w.ctr, w.sphys}, stride_grid(w.capacity));
        copyback.add({w.c, w.alt, w.ctr, w.sphys, w.spos}, stride_grid(w.capacity));
    }
    launch_k<KResortGather<P>>(L, "resort_gather", gather.args, gather.blocks);
    launch_k<KResortCopyback<P>>(L, "resort_copyback", copyback.args, copyback.blocks);
}

template <class P>
void launch_contact_walk(const Launch& L, Views<P> ws, Ticks ts) {
    if (!ts[0].contact) return;
    Batch<KContactWalk<P>> b;
    // One block per record tile rather than the usual capped grid: one
    // wave's staging waits overlap another's screening.
    for (size_t i = 0; i < ws.size(); ++i)
        b.add({ws[i].ctr, ws[i].hg, ws[i].rec, ws[i].bstart, ws[i].fs, ts[i].k_contact}, tile_grid(ws[i].capacity));
    // Five waves per SIMD, as the LDS allows (28 KB per block).
    launch_k<KContactWalk<P>, 5>(L, "contact_walk", b.args, b.blocks);
}

template <class P>
void launch_feed_walk(const Launch& L, Views<P> ws, Ticks ts, const std::vector<char>& crowded) {
    Batch<KFeedWalk<P>> sparse, dense;
    for (size_t i = 0; i < ws.size(); ++i) {
        const WorldDev<P>& w = ws[i];
        const TickParams& t = ts[i];
        const typename KFeedWalk<P>::Args a{w.ctr, w.hg, w.rec, w.bstart, w.g[ASTRO_FIELD_IRRADIANCE],
                                            t.ldx, t.ldy, t.ldz, t.ambient_irr, t.occlusion_exact, w.dirs};
        // crowded: a block per BLOCK / FEED_GROUP cells, grid-stride past 16384 blocks
        if (crowded[i])
            dense.add(a, blocks_for_xcd(std::min((w.capacity + BLOCK / FEED_GROUP - 1) / (BLOCK / FEED_GROUP), 16384)));
        else
            sparse.add(a, cell_grid(w.capacity));
    }
    if (!dense.args.empty()) launch_k<KFeedWalkRows<P>>(L, "feed_walk", dense.args, dense.blocks);
    if (!sparse.args.empty()) launch_k<KFeedWalk<P>>(L, "feed_walk", sparse.args, sparse.blocks);
}

#define INSTANTIATE(P)                                                         \
    template void launch_hash_build<P>(const Launch&, Views<P>, bool, bool);   \
    template void launch_contact_walk<P>(const Launch&, Views<P>, Ticks);      \
    template void launch_feed_walk<P>(const Launch&, Views<P>, Ticks, const std::vector<char>&); \
    template void launch_resort<P>(const Launch&, Views<P>);
INSTANTIATE(double)
INSTANTIATE(float)
#undef INSTANTIATE

}  // namespace astro_dev
