// kernels_tau.hip -- Taumoeba, the predator (src/sim/predation.cu,
// src/sim/predation.cuh).
//
// A second, small population (at most 65 536) on the cells' spatial hash:
//   * k_tau_hunt: each live predator walks the 27 buckets around it in the
//     hash's fixed order, counts live cells (its run-and-tumble signal) and
//     claims the first overlapping one with an atomicMin of its id, then
//     tumbles or runs, crawls (the cells' exact OU step under a constant
//     thrust) and digests on the biology clock;
//   * k_tau_resolve: the lowest-id claimant engulfs its prey;
//   * k_tau_unclaim: puts the touched claim words back to all-ones, so the
//     claim array never needs a clearing pass (or a memset in the graph);
//   * k_tau_nitrogen: one uniform per live predator, drawn whatever happens,
//     and the N2 hazard;
//   * k_tau_life: division (exclusive prefix over the divide flags, as the
//     cells') and, when enabled, stable compaction -- one block, since the
//     store is small, so counts stay on the device and the tick never waits.
// Claims are keyed by the prey's reference slot, as the reference's are; the
// predator also keeps the prey's sorted hash record to find its entry.
#include "kernels.h"

namespace astro_dev {
namespace {

constexpr int TAU_LIFE_THREADS = 1024;

__device__ __forceinline__ bool alive_flags(uint32_t f) {
    return (f & ASTRO_FLAG_OCCUPIED) && (f & ASTRO_FLAG_ALIVE);
}

// gaussian3 then normalize (src/core/rng.cuh:111-117, src/core/vec.cuh:32-35).
__device__ __forceinline__ void random_heading(Pcg& r, double& x, double& y, double& z) {
    double a, b, c, d;
    r.gaussian_pair(a, b);
    r.gaussian_pair(c, d);
    const double l = sqrt(a * a + b * b + c * c);
    if (l > 0.0) {
        const double inv = 1.0 / l;
        x = a * inv;
        y = b * inv;
        z = c * inv;
    } else {
        x = 0.0;
        y = 0.0;
        z = 1.0;
    }
}

// tau_drag (src/sim/predation.cuh:24-26).
__device__ __forceinline__ double tau_drag(double t) {
    return 6.0 * ASTRO_PI * water_viscosity(t) * ASTRO_TAU_RADIUS;
}

__global__ __launch_bounds__(BLOCK) void k_tau_spawn(TauDev t, int32_t count, int32_t base, uint64_t first_id,
                                                     uint64_t seed, double cw, double ch, double cd) {
    const int i = static_cast<int>(blockIdx.x) * BLOCK + static_cast<int>(threadIdx.x);
    if (i >= count) return;
    const int s = base + i;
    const uint64_t id = first_id + static_cast<uint64_t>(i);
    // the cells' recipe on a double-mixed seed: disjoint streams
    Pcg r = cell_rng(cell_rng_init(splitmix64(seed), id), id);
    const double a = ASTRO_TAU_RADIUS;
    t.x[s] = r.uniform(-0.5 * cw + a, 0.5 * cw - a);
    t.y[s] = r.uniform(-0.5 * ch + a, 0.5 * ch - a);
    t.z[s] = r.uniform(-0.5 * cd + a, 0.5 * cd - a);
    t.vx[s] = 0.0;
    t.vy[s] = 0.0;
    t.vz[s] = 0.0;
    double hx, hy, hz;
    random_heading(r, hx, hy, hz);
    t.dir_x[s] = static_cast<float>(hx);
    t.dir_y[s] = static_cast<float>(hy);
    t.dir_z[s] = static_cast<float>(hz);
    t.id[s] = id;
    t.flags[s] = ASTRO_FLAG_OCCUPIED | ASTRO_FLAG_ALIVE;
    t.biomass[s] = ASTRO_TAU_MASS_DRY;
    t.prey_biomass[s] = 0.0;
    t.digest_timer[s] = 0.0f;
    t.tolerance[s] = static_cast<float>(ASTRO_TAU_N2_TOLERANCE_INIT);
    t.generation[s] = 0u;
    t.rng_state[s] = r.state;
    t.density_ema[s] = 0.0f;
    t.run_timer[s] = 0.0f;
    t.target[s] = -1;
}

// hunt_kernel (src/sim/predation.cu:78-159).
template <class P>
__global__ __launch_bounds__(BLOCK) void k_tau_hunt(TauDev t, WorldDev<P> w, TauTick tt) {
    const int n = t.ctr->count;
    const int stride = static_cast<int>(gridDim.x) * BLOCK;
    for (int i = static_cast<int>(blockIdx.x) * BLOCK + static_cast<int>(threadIdx.x); i < n; i += stride) {
        if (!alive_flags(t.flags[i])) continue;
        double px = t.x[i], py = t.y[i], pz = t.z[i];
        const bool digesting = t.digest_timer[i] > 0.0f;

        // live-cell count and the first overlapping live cell, one walk in
        // the hash's order (z, y, x, then sorted position in the bucket)
        int prey_count = 0;
        int32_t prey = -1;
        const double reach = ASTRO_TAU_RADIUS + ASTRO_CELL_RADIUS;
        int gx, gy, gz;
        hash_coords(w.hg, px, py, pz, gx, gy, gz);
        for (int dz = -1; dz <= 1; ++dz) {
            const int cz = gz + dz;
            if (cz < 0 || cz >= w.hg.nz) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                const int cy = gy + dy;
                if (cy < 0 || cy >= w.hg.ny) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    const int cx = gx + dx;
                    if (cx < 0 || cx >= w.hg.nx) continue;
                    const int b = (cz * w.hg.ny + cy) * w.hg.nx + cx;
                    const int e = w.bstart[b + 1];
                    for (int k = w.bstart[b]; k < e; ++k) {
                        const uint32_t p = w.sphys[k];
                        if (!alive_flags(w.c.flags[p])) continue;
                        ++prey_count;
                        if (prey >= 0) continue;
                        const double ex = px - static_cast<double>(w.c.x[p]);
                        const double ey = py - static_cast<double>(w.c.y[p]);
                        const double ez = pz - static_cast<double>(w.c.z[p]);
                        if (ex * ex + ey * ey + ez * ez < reach * reach) prey = k;
                    }
                }
            }
        }
        if (!digesting && prey >= 0) {
            const uint32_t slot = w.vals_sorted[prey];
            t.target[i] = static_cast<int32_t>(slot);
            t.prey_rec[i] = prey;
            atomicMin(&t.claim[slot], static_cast<unsigned long long>(t.id[i]));
        } else {
            t.target[i] = -1;
        }

        // run-and-tumble on the prey count (tau_should_tumble,
        // src/sim/predation.cuh:36-41), then the crawl
        Pcg r = cell_rng(t.rng_state[i], t.id[i]);
        float ema = t.density_ema[i];
        float run = t.run_timer[i];
        double hx = t.dir_x[i], hy = t.dir_y[i], hz = t.dir_z[i];
        const double alpha = -expm1(-tt.dt / ASTRO_TAXIS_MEMORY_TIME);
        const double delta = static_cast<double>(static_cast<float>(prey_count)) - static_cast<double>(ema);
        ema = static_cast<float>(ema + alpha * delta);
        if (!(delta > 0.0 && run < ASTRO_TAXIS_RUN_MAX)) {
            random_heading(r, hx, hy, hz);
            run = 0.0f;
        } else {
            run += static_cast<float>(tt.dt);
        }
        t.density_ema[i] = ema;
        t.run_timer[i] = run;
        t.dir_x[i] = static_cast<float>(hx);
        t.dir_y[i] = static_cast<float>(hy);
        t.dir_z[i] = static_cast<float>(hz);

        // integrate_cell (src/sim/integrator.cuh:154-166) with the thrust along
        // the heading; neutrally buoyant, so no gravity
        double vx = t.vx[i], vy = t.vy[i], vz = t.vz[i];
        const double gamma = tau_drag(tt.ambient_temp);
        const Ou oc = ou_coefficients(ASTRO_TAU_MASS, gamma, tt.ambient_temp, tt.dt);
        double z1, z2, z3, z4, z5, z6;
        r.gaussian_pair(z1, z2);
        r.gaussian_pair(z3, z4);
        r.gaussian_pair(z5, z6);
        const bool noise = tt.thermal_noise != 0;
        ou_advance_axis(px, vx, hx * ASTRO_TAU_CRAWL_THRUST, gamma, oc, tt.dt, noise, z1, z2);
        ou_advance_axis(py, vy, hy * ASTRO_TAU_CRAWL_THRUST, gamma, oc, tt.dt, noise, z3, z4);
        ou_advance_axis(pz, vz, hz * ASTRO_TAU_CRAWL_THRUST, gamma, oc, tt.dt, noise, z5, z6);
        t.rng_state[i] = r.state;
        const double a = ASTRO_TAU_RADIUS;
        apply_boundary_axis(px, vx, 0.5 * tt.chamber_w, a, tt.boundary_x);
        apply_boundary_axis(py, vy, 0.5 * tt.chamber_h, a, tt.boundary_y);
        apply_boundary_axis(pz, vz, 0.5 * tt.chamber_d, a, ASTRO_BOUNDARY_REFLECTING);
        t.x[i] = px;
        t.y[i] = py;
        t.z[i] = pz;
        t.vx[i] = vx;
        t.vy[i] = vy;
        t.vz[i] = vz;

        // digestion on the biology clock
        if (digesting) {
            float d = t.digest_timer[i] - static_cast<float>(tt.dt_bio);
            if (d <= 0.0f) {
                t.biomass[i] += t.prey_biomass[i] * ASTRO_TAU_BIOMASS_YIELD;
                t.prey_biomass[i] = 0.0;
                d = 0.0f;
            }
            t.digest_timer[i] = d;
        }
    }
}

// resolve_kernel (src/sim/predation.cu:163-180): the winning claimant engulfs.
template <class P>
__global__ __launch_bounds__(BLOCK) void k_tau_resolve(TauDev t, WorldDev<P> w, int disposition) {
    const int n = t.ctr->count;
    const int stride = static_cast<int>(gridDim.x) * BLOCK;
    for (int i = static_cast<int>(blockIdx.x) * BLOCK + static_cast<int>(threadIdx.x); i < n; i += stride) {
        const int32_t slot = t.target[i];
        if (slot < 0 || slot >= t.cell_capacity) continue;
        if (t.claim[slot] != static_cast<unsigned long long>(t.id[i])) continue;   // lost the claim
        const uint32_t p = w.sphys[t.prey_rec[i]];
        const uint32_t cf = w.c.flags[p];
        if (!alive_flags(cf)) continue;
        w.c.flags[p] = cf & ~ASTRO_FLAG_ALIVE;
        w.c.death_cause[p] = ASTRO_DEATH_PREDATED;
        w.c.emit_power[p] = 0.0f;
        w.c.energy[p] = disposition == ASTRO_STORE_RETAIN ? w.c.energy[p] : 0.0;
        t.digest_timer[i] = static_cast<float>(ASTRO_TAU_DIGEST_TIME);
        t.prey_biomass[i] = w.c.biomass[p];
    }
}

// Back to the all-ones sentinel for every word a predator may have touched.
__global__ __launch_bounds__(BLOCK) void k_tau_unclaim(TauDev t) {
    const int n = t.ctr->count;
    const int stride = static_cast<int>(gridDim.x) * BLOCK;
    for (int i = static_cast<int>(blockIdx.x) * BLOCK + static_cast<int>(threadIdx.x); i < n; i += stride) {
        const int32_t k = t.target[i];
        if (k >= 0 && k < t.cell_capacity) t.claim[k] = ~0ull;
    }
}

// nitrogen_kernel (src/sim/predation.cu:187-202) with tau_n2_hazard and
// tau_n2_death_prob (src/sim/predation.cuh:50-63).
__global__ __launch_bounds__(BLOCK) void k_tau_nitrogen(TauDev t, Grid n2, int n2_on, double dt_bio) {
    const int n = t.ctr->count;
    const int stride = static_cast<int>(gridDim.x) * BLOCK;
    for (int i = static_cast<int>(blockIdx.x) * BLOCK + static_cast<int>(threadIdx.x); i < n; i += stride) {
        const uint32_t f = t.flags[i];
        if (!alive_flags(f)) continue;
        Pcg r = cell_rng(t.rng_state[i], t.id[i]);
        const double u = r.uniform01d();   // always drawn
        t.rng_state[i] = r.state;
        const double n2_local = n2_on ? static_cast<double>(grid_sample(n2, static_cast<float>(t.x[i]),
                                                                         static_cast<float>(t.y[i]),
                                                                         static_cast<float>(t.z[i])))
                                      : 0.0;
        const double threshold =
            ASTRO_TAU_N2_LETHAL_CONC * (1.0 + static_cast<double>(t.tolerance[i]) * ASTRO_TAU_N2_TOLERANCE_K);
        const double h = n2_local - threshold;
        const double hazard = h > 0.0 ? h : 0.0;
        if (u < -expm1(-hazard * ASTRO_TAU_N2_HAZARD_RATE * dt_bio)) t.flags[i] = f & ~ASTRO_FLAG_ALIVE;
    }
}

// Exclusive prefix of flag[0, n) into out[0, n) by one block: each thread
// owns one contiguous chunk; returns the total.
__device__ int block_exclusive_scan(const int32_t* flag, int32_t* out, int n) {
    __shared__ int part[TAU_LIFE_THREADS];
    const int tid = static_cast<int>(threadIdx.x);
    const int per = (n + TAU_LIFE_THREADS - 1) / TAU_LIFE_THREADS;
    const int lo = tid * per, hi = min(n, lo + per);
    int sum = 0;
    for (int i = lo; i < hi; ++i) sum += flag[i];
    part[tid] = sum;
    __syncthreads();
    for (int off = 1; off < TAU_LIFE_THREADS; off <<= 1) {
        const int v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - sum;
    for (int i = lo; i < hi; ++i) {
        out[i] = run;
        run += flag[i];
    }
    const int total = part[TAU_LIFE_THREADS - 1];
    __syncthreads();
    return total;
}

// One array of the compaction: survivors to [0, live) through the scratch.
template <class T>
__device__ void compact_array(T* a, const int32_t* src, unsigned long long* scratch, int live) {
    T* sc = reinterpret_cast<T*>(scratch);
    for (int k = static_cast<int>(threadIdx.x); k < live; k += TAU_LIFE_THREADS) sc[k] = a[src[k]];
    __syncthreads();
    for (int k = static_cast<int>(threadIdx.x); k < live; k += TAU_LIFE_THREADS) a[k] = sc[k];
    __syncthreads();
}

// Division (tau_mark_kernel + scan + tau_divide_kernel,
// src/sim/predation.cu:208-280, :463-481) and compaction
// (taumoeba_store_compact, :492-527), one block.
__global__ __launch_bounds__(TAU_LIFE_THREADS) void k_tau_life(TauDev t, TauTick tt) {
    __shared__ int s_deads;
    const int tid = static_cast<int>(threadIdx.x);
    const int n = t.ctr->count;
    if (tid == 0) s_deads = 0;
    __syncthreads();
    int deads = 0;
    for (int i = tid; i < n; i += TAU_LIFE_THREADS) {
        const uint32_t f = t.flags[i];
        const bool occ = (f & ASTRO_FLAG_OCCUPIED) != 0u, alive = (f & ASTRO_FLAG_ALIVE) != 0u;
        t.mark[i] = (occ && alive && t.biomass[i] >= ASTRO_TAU_DIVIDE_BIOMASS) ? 1 : 0;
        if (occ && !alive) ++deads;
    }
    if (deads) atomicAdd(&s_deads, deads);
    const int births_all = block_exclusive_scan(t.mark, t.prefix, n);
    const int room = t.capacity - n;
    const int births = births_all < room ? births_all : (room > 0 ? room : 0);
    const uint64_t first_id = t.ctr->next_id;
    if (births > 0) {
        for (int i = tid; i < n; i += TAU_LIFE_THREADS) {
            if (!t.mark[i]) continue;
            const int s = n + t.prefix[i];
            if (s >= t.capacity) continue;   // store full: skip, do not wrap
            const uint64_t did = first_id + static_cast<uint64_t>(t.prefix[i]);
            const double half = 0.5 * t.biomass[i];
            t.biomass[i] = half;
            t.biomass[s] = half;
            Pcg dr = cell_rng(pcg_split(t.rng_state[i], did), did);
            double mutation, unused;
            dr.gaussian_pair(mutation, unused);
            const double tol = static_cast<double>(t.tolerance[i]) + ASTRO_TAU_MUTATION_SIGMA * mutation;
            t.tolerance[s] = static_cast<float>(tol < 0.0 ? 0.0 : (tol > 1.0 ? 1.0 : tol));
            t.rng_state[s] = dr.state;
            t.generation[s] = t.generation[i] + 1u;
            double ax, ay, az;
            division_axis(did, ax, ay, az);
            const double a = ASTRO_TAU_RADIUS;
            const double hx = 0.5 * tt.chamber_w - a, hy = 0.5 * tt.chamber_h - a, hz = 0.5 * tt.chamber_d - a;
            auto cl = [](double v, double h) { return v < -h ? -h : (v > h ? h : v); };
            const double px = t.x[i], py = t.y[i], pz = t.z[i];
            t.x[s] = cl(px + a * ax, hx);
            t.y[s] = cl(py + a * ay, hy);
            t.z[s] = cl(pz + a * az, hz);
            t.x[i] = cl(px - a * ax, hx);
            t.y[i] = cl(py - a * ay, hy);
            t.z[i] = cl(pz - a * az, hz);
            t.id[s] = did;
            t.flags[s] = t.flags[i];
            t.vx[s] = t.vx[i];
            t.vy[s] = t.vy[i];
            t.vz[s] = t.vz[i];
            t.dir_x[s] = t.dir_x[i];
            t.dir_y[s] = t.dir_y[i];
            t.dir_z[s] = t.dir_z[i];
            t.prey_biomass[s] = 0.0;
            t.digest_timer[s] = 0.0f;
            t.density_ema[s] = 0.0f;
            t.run_timer[s] = 0.0f;
            t.target[s] = -1;
        }
    }
    __syncthreads();
    const int cnt = n + births;
    const int dead_slots = s_deads;
    if (tid == 0) {
        t.ctr->births = births;
        t.ctr->deads = dead_slots;
        t.ctr->count = cnt;
        t.ctr->next_id = first_id + static_cast<uint64_t>(births);
    }
    if (!tt.compaction || dead_slots == 0 || cnt == 0) return;
    // stable compaction of the survivors
    for (int i = tid; i < cnt; i += TAU_LIFE_THREADS) t.mark[i] = alive_flags(t.flags[i]) ? 1 : 0;
    __syncthreads();
    const int live = block_exclusive_scan(t.mark, t.prefix, cnt);
    if (live >= cnt) return;
    for (int i = tid; i < cnt; i += TAU_LIFE_THREADS)
        if (t.mark[i]) t.src[t.prefix[i]] = i;
    __syncthreads();
    compact_array(t.id, t.src, t.scratch, live);
    compact_array(t.flags, t.src, t.scratch, live);
    compact_array(t.x, t.src, t.scratch, live);
    compact_array(t.y, t.src, t.scratch, live);
    compact_array(t.z, t.src, t.scratch, live);
    compact_array(t.vx, t.src, t.scratch, live);
    compact_array(t.vy, t.src, t.scratch, live);
    compact_array(t.vz, t.src, t.scratch, live);
    compact_array(t.dir_x, t.src, t.scratch, live);
    compact_array(t.dir_y, t.src, t.scratch, live);
    compact_array(t.dir_z, t.src, t.scratch, live);
    compact_array(t.biomass, t.src, t.scratch, live);
    compact_array(t.prey_biomass, t.src, t.scratch, live);
    compact_array(t.digest_timer, t.src, t.scratch, live);
    compact_array(t.tolerance, t.src, t.scratch, live);
    compact_array(t.generation, t.src, t.scratch, live);
    compact_array(t.rng_state, t.src, t.scratch, live);
    compact_array(t.density_ema, t.src, t.scratch, live);
    compact_array(t.run_timer, t.src, t.scratch, live);
    compact_array(t.target, t.src, t.scratch, live);
    if (tid == 0) t.ctr->count = live;
}

// stats_taumoeba_kernel (src/sim/stats.cu:65-74).
__global__ __launch_bounds__(BLOCK) void k_tau_stats(TauDev t, unsigned long long* acc) {
    const int n = t.ctr->count;
    const int stride = static_cast<int>(gridDim.x) * BLOCK;
    for (int i = static_cast<int>(blockIdx.x) * BLOCK + static_cast<int>(threadIdx.x); i < n; i += stride) {
        if (!alive_flags(t.flags[i])) continue;
        atomicAdd(&acc[SW_N_TAU], 1ull);
        deposit_add(&acc[SW_TAU_TOL], static_cast<double>(t.tolerance[i]), ASTRO_DEPOSIT_SCALE_STATS);
    }
}

__global__ __launch_bounds__(BLOCK) void k_fill_u64(unsigned long long* v, size_t n, unsigned long long value) {
    const size_t stride = static_cast<size_t>(gridDim.x) * BLOCK;
    for (size_t i = static_cast<size_t>(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += stride) v[i] = value;
}

unsigned tau_blocks(int capacity) {
    const int b = (capacity + BLOCK - 1) / BLOCK;
    return static_cast<unsigned>(b < 1 ? 1 : b);
}

}  // namespace

template <class P>
void launch_predation(const Launch& L, const WorldDev<P>& w, const TauDev& t, const TauTick& tt) {
    const unsigned blocks = tau_blocks(t.capacity);
    k_tau_hunt<P><<<blocks, BLOCK, 0, L.stream>>>(t, w, tt);
    k_tau_resolve<P><<<blocks, BLOCK, 0, L.stream>>>(t, w, tt.store_disposition);
    k_tau_unclaim<<<blocks, BLOCK, 0, L.stream>>>(t);
    k_tau_nitrogen<<<blocks, BLOCK, 0, L.stream>>>(t, w.g[ASTRO_FIELD_N2], w.on[ASTRO_FIELD_N2] ? 1 : 0, tt.dt_bio);
    k_tau_life<<<1, TAU_LIFE_THREADS, 0, L.stream>>>(t, tt);
}

void launch_tau_spawn(const Launch& L, const TauDev& t, int32_t count, int32_t base, uint64_t first_id,
                      uint64_t seed, double chamber_w, double chamber_h, double chamber_d) {
    if (count <= 0) return;
    k_tau_spawn<<<tau_blocks(count), BLOCK, 0, L.stream>>>(t, count, base, first_id, seed, chamber_w, chamber_h,
                                                           chamber_d);
}

void launch_fill_u64(const Launch& L, unsigned long long* v, size_t n, unsigned long long value) {
    if (n == 0) return;
    const size_t want = (n + BLOCK - 1) / BLOCK;
    k_fill_u64<<<static_cast<unsigned>(want < 4096 ? want : 4096), BLOCK, 0, L.stream>>>(v, n, value);
}

void launch_tau_stats(const Launch& L, const TauDev& t, unsigned long long* stats) {
    k_tau_stats<<<tau_blocks(t.capacity), BLOCK, 0, L.stream>>>(t, stats);
}

template void launch_predation<double>(const Launch&, const WorldDev<double>&, const TauDev&, const TauTick&);
template void launch_predation<float>(const Launch&, const WorldDev<float>&, const TauDev&, const TauTick&);

}  // namespace astro_dev
