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
#include "kernels.h"

namespace astro_dev {

namespace {

template <class P>
__global__ __launch_bounds__(BLOCK) void k_hash_key(Cells<P> c, Counters* ctr, HashGeo h,
                                                    uint32_t* keys, uint32_t* rin, uint32_t* bcount,
                                                    int count_tick) {
    const int n = ctr->count;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (count_tick && i == 0) ctr->cell_steps += n;
    if (i >= n) return;
    const uint32_t f = c.flags[i];
    const uint32_t key = (f & ASTRO_FLAG_OCCUPIED)
        ? hash_key(h, static_cast<double>(c.x[i]), static_cast<double>(c.y[i]), static_cast<double>(c.z[i]))
        : static_cast<uint32_t>(h.nb);
    keys[i] = key;
    rin[i] = atomicAdd(&bcount[key], 1u);
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
                                                        const uint32_t* rin, const int32_t* bstart, Rec<P>* trec) {
    const int n = ctr->count;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t f = c.flags[i];
    Rec<P> r;
    r.x = c.x[i];
    r.y = c.y[i];
    r.z = c.z[i];
    r.meta = static_cast<uint32_t>(i) | ((f & ASTRO_FLAG_OCCUPIED) ? 0u : REC_UNOCC) |
             ((f & ASTRO_FLAG_ALIVE) ? REC_ALIVE : 0u);
    if constexpr (sizeof(Rec<P>) == 32) r._pad = 0u;
    trec[bstart[keys[i]] + static_cast<int32_t>(rin[i])] = r;
}

template <class P>
__global__ __launch_bounds__(BLOCK) void k_hash_rank(const Counters* ctr, HashGeo h, const Rec<P>* trec,
                                                     const int32_t* bstart, Rec<P>* rec, uint32_t* vals_sorted,
                                                     int32_t* spos) {
    const int n = ctr->count;
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= n) return;
    const Rec<P> r = trec[p];
    const uint32_t slot = r.meta & REC_SLOT_MASK;
    const uint32_t key = (r.meta & REC_UNOCC)
        ? static_cast<uint32_t>(h.nb)
        : hash_key(h, static_cast<double>(r.x), static_cast<double>(r.y), static_cast<double>(r.z));
    const int32_t s = bstart[key], e = bstart[key + 1];
    int32_t rank = 0;
    for (int32_t q = s; q < e; ++q) rank += ((trec[q].meta & REC_SLOT_MASK) < slot) ? 1 : 0;
    const int32_t dst = s + rank;
    rec[dst] = r;
    vals_sorted[dst] = slot;
    spos[slot] = dst;
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
    for (int dz = -1; dz <= 1; ++dz) {
        const int cz = gz + dz;
        if (cz < 0 || cz >= h.nz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int cy = gy + dy;
            if (cy < 0 || cy >= h.ny) continue;
            const int row = (cz * h.ny + cy) * h.nx;
            const int32_t s = bstart[row + xlo], e = bstart[row + xhi + 1];
            for (int32_t q = s; q < e; ++q) fn(rec[q]);
        }
    }
}

// forces_kernel's contact half (src/sim/integrator.cu:60-67), one thread per
// sorted record, starting from the buoyancy + recoil force already in fs.
template <class P>
__global__ __launch_bounds__(BLOCK) void k_forces_walk(const Counters* ctr, HashGeo h, const Rec<P>* rec,
                                                       const int32_t* bstart, Force* fs, double k_contact) {
    const int n = ctr->count;
    const int nl = gridDim.x;
    const int b = xcd_tile(blockIdx.x, nl);
    const int k = b * BLOCK + threadIdx.x;
    if (k >= n) return;
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
}

// feed_kernel's irradiance half (src/sim/emission.cu:81-105): far-field sample
// times the exact near-field shadow product over upstream live neighbours.
template <class P>
__global__ __launch_bounds__(BLOCK) void k_feed_walk(const Counters* ctr, HashGeo h, const Rec<P>* rec,
                                                     const int32_t* bstart, Grid irr, double ldx, double ldy,
                                                     double ldz, float ambient, int occl, double* dirs) {
    const int n = ctr->count;
    const int nl = gridDim.x;
    const int b = xcd_tile(blockIdx.x, nl);
    const int k = b * BLOCK + threadIdx.x;
    if (k >= n) return;
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
}

inline int grid_for(int n) { return (n + BLOCK - 1) / BLOCK; }

}  // namespace

template <class P>
void launch_hash_build(const Launch& L, const WorldDev<P>& w, bool count_tick) {
    const int g = grid_for(w.capacity);
    if (L.prof) L.prof->begin("hash_key");
    k_hash_key<P><<<g, BLOCK, 0, L.stream>>>(w.c, w.ctr, w.hg, w.keys, w.rin, w.bcount, count_tick ? 1 : 0);
    if (L.prof) L.prof->end();
    const int nbp1 = w.hg.nb + 1;
    if (L.prof) L.prof->begin("hash_bucket_scan");
    k_bucket_scan<<<scan_tiles_for(nbp1), BLOCK, 0, L.stream>>>(w.bcount, w.bstart, nbp1, w.sc);
    if (L.prof) L.prof->end();
    if (L.prof) L.prof->begin("hash_scatter");
    k_hash_scatter<P><<<g, BLOCK, 0, L.stream>>>(w.c, w.ctr, w.keys, w.rin, w.bstart, w.trec);
    if (L.prof) L.prof->end();
    if (L.prof) L.prof->begin("hash_rank");
    k_hash_rank<P><<<g, BLOCK, 0, L.stream>>>(w.ctr, w.hg, w.trec, w.bstart, w.rec, w.vals_sorted, w.spos);
    if (L.prof) L.prof->end();
}

template <class P>
void launch_forces_walk(const Launch& L, const WorldDev<P>& w, const TickParams& t) {
    if (!t.contact) return;
    const int g = blocks_for_xcd(grid_for(w.capacity));
    if (L.prof) L.prof->begin("forces_walk");
    k_forces_walk<P><<<g, BLOCK, 0, L.stream>>>(w.ctr, w.hg, w.rec, w.bstart, w.fs, t.k_contact);
    if (L.prof) L.prof->end();
}

template <class P>
void launch_feed_walk(const Launch& L, const WorldDev<P>& w, const TickParams& t) {
    const int g = blocks_for_xcd(grid_for(w.capacity));
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

}  // namespace astro_dev
