// kernels_hash.hip -- spatial-hash build and the two neighbour walks.
//
// The reference builds the hash with a CUB stable radix sort of (bucket, slot)
// pairs plus an adjacent-key range pass (src/sim/hash.cu:106-133). Here the
// key *is* the sort digit, so the build is a counting sort over the bucket
// index (a one-digit LSD radix sort):
//   1. key:     bucket of each slot; arrival rank = atomicAdd on its count.
//   2. scan:    dense exclusive prefix of the counts (look-back scan); the
//               counts are zeroed in the same pass for the next build.
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
__global__ __launch_bounds__(BLOCK) void k_hash_key(Cells<P> c, Counters* ctr, HashGeo h, uint32_t* keys,
                                                    uint32_t* rin, uint32_t* bcount, int32_t* dflag, int32_t* drop,
                                                    int cap, int count_tick, int resort_period) {
    const int np = ctr->nphys;
    const int gid = blockIdx.x * BLOCK + threadIdx.x, stride = gridDim.x * BLOCK;
    if (count_tick) {
        if (gid == 0) {
            const int n = ctr->count;
            ctr->cell_steps += n;
            const uint32_t t = ctr->tick_ctr;
            ctr->resort = (t % static_cast<uint32_t>(resort_period) == 0u || np > cap) ? 1 : 0;
            ctr->tick_ctr = t + 1u;
        }
        // The lifecycle's per-slot flags start each tick cleared.
        for (int q = gid; q < cap; q += stride) {
            dflag[q] = 0;
            drop[q] = 0;
        }
    }
    for (int p = gid; p < np; p += stride) {
        if (c.lslot[p] == NO_SLOT) continue;
        const uint32_t f = c.flags[p];
        const uint32_t key = (f & ASTRO_FLAG_OCCUPIED)
            ? hash_key(h, static_cast<double>(c.x[p]), static_cast<double>(c.y[p]), static_cast<double>(c.z[p]))
            : static_cast<uint32_t>(h.nb);
        keys[p] = key;
        rin[p] = atomicAdd(&bcount[key], 1u);
    }
}

__global__ __launch_bounds__(BLOCK) void k_bucket_scan(uint32_t* bcount, int32_t* bstart, int nbp1, ScanCtl sc) {
    int32_t* counts = reinterpret_cast<int32_t*>(bcount);
    scan_exclusive(
        nbp1,
        [&](int i) {
            const int4 v = load4_i32(counts, i, nbp1);
            if (v.x | v.y | v.z | v.w) store4_i32(counts, i, nbp1, int4{0, 0, 0, 0});
            return v;
        },
        [&](int i, int4 ex) { store4_i32(bstart, i, nbp1, ex); }, sc, bstart + nbp1);
}

template <class P>
__global__ __launch_bounds__(BLOCK) void k_hash_scatter(Cells<P> c, const Counters* ctr, const uint32_t* keys,
                                                        const uint32_t* rin, const int32_t* bstart, Rec<P>* trec,
                                                        uint32_t* tphys) {
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
    for (int p = blockIdx.x * BLOCK + threadIdx.x; p < np; p += gridDim.x * BLOCK) body(p);
}

// Stable order: ascending reference slot inside each bucket.
template <class P>
__global__ __launch_bounds__(BLOCK) void k_hash_rank(const Counters* ctr, HashGeo h, const Rec<P>* trec,
                                                     const uint32_t* tphys, const int32_t* bstart, Rec<P>* rec,
                                                     uint32_t* vals_sorted, uint32_t* sphys, int32_t* spos) {
    const int n = ctr->count;
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
    for (int k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += gridDim.x * BLOCK) body(k);
}

// Re-sort: move every live entry to its sorted position (drops tombstones),
// then copy back. Runs only on ticks the key kernel marked.
template <class P>
__global__ __launch_bounds__(BLOCK) void k_resort_gather(Cells<P> a, Cells<P> b, const Counters* ctr,
                                                         const uint32_t* sphys) {
    if (!ctr->resort) return;
    const int n = ctr->count;
    const auto body = [&](const int k) {
    copy_cell(a, static_cast<int>(sphys[k]), b, k);
    };
    for (int k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += gridDim.x * BLOCK) body(k);
}

template <class P>
__global__ __launch_bounds__(BLOCK) void k_resort_copyback(Cells<P> a, Cells<P> b, const Counters* ctr,
                                                           uint32_t* sphys, int32_t* spos) {
    if (!ctr->resort) return;
    const int n = ctr->count;
    const auto body = [&](const int k) {
    copy_cell(b, k, a, k);
    sphys[k] = static_cast<uint32_t>(k);
    spos[k] = k;
    };
    for (int k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += gridDim.x * BLOCK) body(k);
}

__global__ void k_resort_finalize(Counters* ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && ctr->resort) {
        ctr->nphys = ctr->count;
        ctr->resort = 0;
    }
}

// Walks the 27 buckets around (px,py,pz) in the reference order (z, then y,
// then x, then sorted position, src/sim/hash.cuh:88-111). The three x-adjacent
// buckets of one (y,z) row are contiguous in the sorted array, so each row is
// a single range [bstart[x_lo], bstart[x_hi + 1]).
template <class P, class Fn>
__device__ __forceinline__ void walk_rows(const HashGeo& h, const int32_t* bstart, const Rec<P>* rec,
                                          double px, double py, double pz, Fn&& fn) {
    int gx, gy, gz;
    hash_coords(h, px, py, pz, gx, gy, gz);
    const int xlo = gx > 0 ? gx - 1 : 0;
    const int xhi = gx < h.nx - 1 ? gx + 1 : h.nx - 1;
    // All nine row ranges, then each row's first record, are requested before
    // any is used: neighbourhoods hold about one record per row, so this turns
    // a chain of ~20 dependent loads into three rounds.
    int32_t s[9], e[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const int cz = gz + r / 3 - 1, cy = gy + r % 3 - 1;
        const bool ok = cz >= 0 && cz < h.nz && cy >= 0 && cy < h.ny;
        const int row = ok ? (cz * h.ny + cy) * h.nx : 0;
        s[r] = ok ? bstart[row + xlo] : 0;
        e[r] = ok ? bstart[row + xhi + 1] : 0;
    }
    Rec<P> first[9];
#pragma unroll
    for (int r = 0; r < 9; ++r)
        if (s[r] < e[r]) first[r] = rec[s[r]];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        if (s[r] >= e[r]) continue;
        fn(first[r]);
        for (int32_t q = s[r] + 1; q < e[r]; ++q) fn(rec[q]);
    }
}

// forces_kernel's contact half (src/sim/integrator.cu:60-67), one thread per
// sorted record, starting from the buoyancy + recoil force already in fs.
template <class P>
__global__ __launch_bounds__(BLOCK) void k_forces_walk(const Counters* ctr, HashGeo h, const Rec<P>* rec,
                                                       const int32_t* bstart, Force* fs, double k_contact) {
    const int n = ctr->count;
    const int lb = xcd_tile(blockIdx.x, gridDim.x);
    const auto body = [&](const int k) {
    const Rec<P> r = rec[k];
    if (r.meta & REC_UNOCC) return;
    const uint32_t self = r.meta & REC_SLOT_MASK;
    const double px = static_cast<double>(r.x), py = static_cast<double>(r.y), pz = static_cast<double>(r.z);
    Force F = fs[k];
    const double two_a = ASTRO_CELL_DIAMETER;
    walk_rows<P>(h, bstart, rec, px, py, pz, [&](const Rec<P>& o) {
        if ((o.meta & REC_SLOT_MASK) == self) return;
        // contact_force (src/sim/integrator.cuh:210-219)
        const double dx = px - static_cast<double>(o.x);
        const double dy = py - static_cast<double>(o.y);
        const double dz = pz - static_cast<double>(o.z);
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 >= two_a * two_a || r2 <= 0.0) {
            F.x += 0.0;
            F.y += 0.0;
            F.z += 0.0;
            return;
        }
        const double rr = sqrt(r2);
        const double overlap = two_a - rr;
        const double s = k_contact * overlap / rr;
        F.x += dx * s;
        F.y += dy * s;
        F.z += dz * s;
    });
    fs[k] = F;
    };
    for (int t = lb; t * BLOCK < n; t += gridDim.x) {
        const int k = t * BLOCK + threadIdx.x;
        if (k < n) body(k);
    }
}

// feed_kernel's irradiance half (src/sim/emission.cu:81-105): far-field sample
// times the exact near-field shadow product over upstream live neighbours.
template <class P>
__global__ __launch_bounds__(BLOCK) void k_feed_walk(const Counters* ctr, HashGeo h, const Rec<P>* rec,
                                                     const int32_t* bstart, Grid irr, double ldx, double ldy,
                                                     double ldz, float ambient, int occl, double* dirs) {
    const int n = ctr->count;
    const int lb = xcd_tile(blockIdx.x, gridDim.x);
    const auto body = [&](const int k) {
    const Rec<P> r = rec[k];
    if ((r.meta & REC_UNOCC) || !(r.meta & REC_ALIVE)) return;
    const uint32_t self = r.meta & REC_SLOT_MASK;
    const double px = static_cast<double>(r.x), py = static_cast<double>(r.y), pz = static_cast<double>(r.z);
    double directional = static_cast<double>(grid_sample(irr, static_cast<float>(r.x), static_cast<float>(r.y),
                                                         static_cast<float>(r.z))) -
                         static_cast<double>(ambient);
    if (directional < 0.0) directional = 0.0;
    if (occl) {
        double open = 1.0;
        walk_rows<P>(h, bstart, rec, px, py, pz, [&](const Rec<P>& o) {
            if ((o.meta & REC_SLOT_MASK) == self || !(o.meta & REC_ALIVE)) return;
            // shadow_fraction (src/sim/emission.cuh:30-37)
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
            open *= (1.0 - frac);
        });
        directional *= open;
    }
    dirs[k] = directional;
    };
    for (int t = lb; t * BLOCK < n; t += gridDim.x) {
        const int k = t * BLOCK + threadIdx.x;
        if (k < n) body(k);
    }
}

inline int grid_for(int n) { return (n + BLOCK - 1) / BLOCK; }

}  // namespace

template <class P>
void launch_hash_build(const Launch& L, const WorldDev<P>& w, bool count_tick) {
    const int gp = stride_grid(w.pcap);
    const int gc = stride_grid(w.capacity);
    if (L.prof) L.prof->begin("hash_key");
    k_hash_key<P><<<gp, BLOCK, 0, L.stream>>>(
        w.c, w.ctr, w.hg, w.keys, w.rin, w.bcount, w.dflag, w.drop, w.capacity, count_tick ? 1 : 0,
        w.resort_period);
    if (L.prof) L.prof->end();
    const int nbp1 = w.hg.nb + 1;
    if (L.prof) L.prof->begin("hash_bucket_scan");
    k_bucket_scan<<<scan_tiles_for(nbp1), BLOCK, 0, L.stream>>>(w.bcount, w.bstart, nbp1, w.sc);
    if (L.prof) L.prof->end();
    if (L.prof) L.prof->begin("hash_scatter");
    k_hash_scatter<P><<<gp, BLOCK, 0, L.stream>>>(w.c, w.ctr, w.keys, w.rin, w.bstart, w.trec, w.tphys);
    if (L.prof) L.prof->end();
    if (L.prof) L.prof->begin("hash_rank");
    k_hash_rank<P><<<gc, BLOCK, 0, L.stream>>>(w.ctr, w.hg, w.trec, w.tphys, w.bstart, w.rec, w.vals_sorted,
                                                w.sphys, w.spos);
    if (L.prof) L.prof->end();
}

template <class P>
void launch_resort(const Launch& L, const WorldDev<P>& w) {
    const int gc = stride_grid(w.capacity);
    if (L.prof) L.prof->begin("resort_gather");
    k_resort_gather<P><<<gc, BLOCK, 0, L.stream>>>(w.c, w.alt, w.ctr, w.sphys);
    if (L.prof) L.prof->end();
    if (L.prof) L.prof->begin("resort_copyback");
    k_resort_copyback<P><<<gc, BLOCK, 0, L.stream>>>(w.c, w.alt, w.ctr, w.sphys, w.spos);
    if (L.prof) L.prof->end();
    if (L.prof) L.prof->begin("resort_finalize");
    k_resort_finalize<<<1, 64, 0, L.stream>>>(w.ctr);
    if (L.prof) L.prof->end();
}

template <class P>
void launch_forces_walk(const Launch& L, const WorldDev<P>& w, const TickParams& t) {
    if (!t.contact) return;
    const int g = blocks_for_xcd(stride_grid(w.capacity));
    if (L.prof) L.prof->begin("forces_walk");
    k_forces_walk<P><<<g, BLOCK, 0, L.stream>>>(w.ctr, w.hg, w.rec, w.bstart, w.fs, t.k_contact);
    if (L.prof) L.prof->end();
}

template <class P>
void launch_feed_walk(const Launch& L, const WorldDev<P>& w, const TickParams& t) {
    const int g = blocks_for_xcd(stride_grid(w.capacity));
    if (L.prof) L.prof->begin("feed_walk");
    k_feed_walk<P><<<g, BLOCK, 0, L.stream>>>(w.ctr, w.hg, w.rec, w.bstart, w.g[ASTRO_FIELD_IRRADIANCE], t.ldx,
                                              t.ldy, t.ldz, t.ambient_irr, t.occlusion_exact, w.dirs);
    if (L.prof) L.prof->end();
}

template void launch_hash_build<double>(const Launch&, const WorldDev<double>&, bool);
template void launch_hash_build<float>(const Launch&, const WorldDev<float>&, bool);
template void launch_forces_walk<double>(const Launch&, const WorldDev<double>&, const TickParams&);
template void launch_forces_walk<float>(const Launch&, const WorldDev<float>&, const TickParams&);
template void launch_feed_walk<double>(const Launch&, const WorldDev<double>&, const TickParams&);
template void launch_feed_walk<float>(const Launch&, const WorldDev<float>&, const TickParams&);
template void launch_resort<double>(const Launch&, const WorldDev<double>&);
template void launch_resort<float>(const Launch&, const WorldDev<float>&);

}  // namespace astro_dev
