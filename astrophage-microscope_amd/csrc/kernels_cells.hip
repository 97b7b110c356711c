// kernels_cells.hip -- per-cell stages: spawn, thermostat, feeding + taxis,
// field sample + body force, the OU integrator, lifecycle and compaction, and
// the per-cell telemetry reduction. One thread per slot; every per-cell array
// is read and written coalesced in slot order.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "cell_ops.h"

#include "kernels.h"

namespace astro_dev {

namespace {

inline int grid_for(int n) { return (n + BLOCK - 1) / BLOCK; }

// 64-lane wave count of a predicate, one atomic per wave. Every lane of the
// wave must call it.
__device__ __forceinline__ void wave_count_add(int32_t* dst, bool pred) {
    const unsigned long long m = __ballot(pred);
    if (m != 0ull && (threadIdx.x & 63) == 0) atomicAdd(dst, static_cast<int32_t>(__popcll(m)));
}

// ---------------------------------------------------------------------------
// spawn_kernel (src/sim/cell_store.cu:42-130); LOG_UNIFORM charge added.
// ---------------------------------------------------------------------------
// One new cell's draws and state, in the reference's draw order.
template <class P, class G>
__device__ __forceinline__ void spawn_cell(const Cells<P>& c, const SpawnArgs& p, int i, int32_t slot, uint64_t id,
                                           G& r) {
    double x = 0.0, y = 0.0;
    if (p.placement == ASTRO_PLACE_GAUSSIAN) {
        double gx, gy;
        r.gaussian_pair(gx, gy);
        x = p.place_x + p.place_radius * gx;
        y = p.place_y + p.place_radius * gy;
    } else if (p.placement == ASTRO_PLACE_DISC) {
        const double rad = p.place_radius * sqrt(r.uniform01d());
        const double ang = r.uniform(0.0, ASTRO_TWO_PI);
        x = p.place_x + rad * cos(ang);
        y = p.place_y + rad * sin(ang);
    } else if (p.placement == ASTRO_PLACE_GRID) {
        const int side = static_cast<int>(ceil(sqrt(static_cast<double>(p.count))));
        const int gx = i % side, gy = i / side;
        x = (static_cast<double>(gx) / side - 0.5) * p.chamber_w * 0.98;
        y = (static_cast<double>(gy) / side - 0.5) * p.chamber_h * 0.98;
    } else {
        x = r.uniform(-0.5 * p.chamber_w, 0.5 * p.chamber_w);
        y = r.uniform(-0.5 * p.chamber_h, 0.5 * p.chamber_h);
    }
    const double a = ASTRO_CELL_RADIUS;
    auto cl = [](double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); };
    x = cl(x, -0.5 * p.chamber_w + a, 0.5 * p.chamber_w - a);
    y = cl(y, -0.5 * p.chamber_h + a, 0.5 * p.chamber_h - a);
    const double z = r.uniform(-0.5 * p.chamber_d + a, 0.5 * p.chamber_d - a);
    double charge;
    if (p.charge_dist == ASTRO_DIST_UNIFORM) {
        charge = r.uniform(p.charge_a, p.charge_b);
    } else if (p.charge_dist == ASTRO_DIST_NORMAL) {
        double z0, z1;
        r.gaussian_pair(z0, z1);
        charge = p.charge_a + p.charge_b * z0;
    } else if (p.charge_dist == ASTRO_DIST_LOG_UNIFORM) {
        charge = exp(r.uniform(log(p.charge_a), log(p.charge_b)));
    } else {
        charge = p.charge_a;
    }
    charge = cl(charge, 0.0, 1.0);
    c.id[slot] = id;
    c.flags[slot] = ASTRO_FLAG_OCCUPIED | ASTRO_FLAG_ALIVE | (p.awake ? ASTRO_FLAG_AWAKE : 0u);
    c.death_cause[slot] = ASTRO_DEATH_NONE;
    c.x[slot] = static_cast<P>(x);
    c.y[slot] = static_cast<P>(y);
    c.z[slot] = static_cast<P>(z);
    c.vx[slot] = P(0);
    c.vy[slot] = P(0);
    c.vz[slot] = P(0);
    c.energy[slot] = charge * ASTRO_CELL_ENERGY_MAX;
    const float tc = static_cast<float>(p.awake ? ASTRO_CELL_TEMP_SETPOINT : ASTRO_AMBIENT_TEMP_DEFAULT);
    c.temp_cell[slot] = tc;
    c.emit_power[slot] = 0.0f;
    double ga, gb, gc, gd;
    r.gaussian_pair(ga, gb);
    r.gaussian_pair(gc, gd);
    const double l = sqrt(ga * ga + gb * gb + gc * gc);
    double dx = 0.0, dy = 0.0, dz = 1.0;
    if (l > 0.0) {
        const double inv = 1.0 / l;
        dx = ga * inv;
        dy = gb * inv;
        dz = gc * inv;
    }
    c.dir_x[slot] = static_cast<float>(dx);
    c.dir_y[slot] = static_cast<float>(dy);
    c.dir_z[slot] = static_cast<float>(dz);
    c.biomass[slot] = ASTRO_CELL_MASS_DRY;
    c.co2_held[slot] = 0.0;
    c.age_s[slot] = 0.0f;
    c.taxis_memory[slot] = 0.0f;
    c.run_timer[slot] = 0.0f;
    c.t_local[slot] = tc;
    c.irradiance[slot] = 0.0f;
    c.co2_local[slot] = 0.0f;
    c.n2_local[slot] = 0.0f;
}

template <class P>
__global__ __launch_bounds__(BLOCK) void k_spawn(Cells<P> c, SpawnArgs p, int32_t base, int32_t lbase,
                                                 uint64_t first_id, uint64_t seed) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= p.count) return;
    const int32_t slot = base + i;
    c.lslot[slot] = static_cast<uint32_t>(lbase + i);
    const uint64_t id = first_id + static_cast<uint64_t>(i);
    if (p.philox) {
        Philox r(seed, RNG_STAGE_SPAWN, 0u, id);
        spawn_cell(c, p, i, slot, id, r);
        c.rng_state[slot] = 0ull;
    } else {
        Pcg r = cell_rng(cell_rng_init(seed, id), id);
        spawn_cell(c, p, i, slot, id, r);
        c.rng_state[slot] = r.state;
    }
}

// ---------------------------------------------------------------------------
// thermal_kernel (src/sim/thermal.cu:19-71). A starving cell also clears the
// alive bit of its sorted neighbour record, so the feed walk later this tick
// sees the same flags the reference's feed_kernel reads.
//
// Two deposit paths, equal bit for bit (mode, below):
//   * ATOMIC: every depositor adds into its voxel's fixed-point accumulator
//     and KApplyAwake flushes it, value += float(from_fixed(total));
//   * DIRECT: the occupied entries of each voxel are counted once per tick
//     (KThermalCount). A cell that is the only entry in its voxel applies its
//     own deposit, value += float(from_fixed(raw)) -- exactly the flush of a
//     single depositor, and no other thread of the pass reads or writes that
//     voxel, since every occupied entry reads its voxel's value -- so only
//     shared voxels use the accumulator. The first substep reads the counts
//     into talone (1 alone, 2 shared), the later ones read talone, and the
//     first apply pass clears the counts again.
// ---------------------------------------------------------------------------
enum : int { THERMAL_ATOMIC = 0, THERMAL_DIRECT_FIRST = 1, THERMAL_DIRECT = 2 };

// The occupied entries counted per temperature voxel: each of them, live or
// dead, reads its voxel's value in the thermal pass. Positions do not change
// during the thermal step, so the counts hold for all its substeps.
template <class P>
struct KThermalCount {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> c;
        const Counters* ctr;
        Grid T;
        uint32_t* tcount;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [c, ctr, T, tcount] = a;
        xcd_tiles(vg, ctr->nphys, [&](const int i, const bool valid) {
            if (!valid) return;
            const uint32_t fl = c.flags[i];
            const float xf = static_cast<float>(c.x[i]), yf = static_cast<float>(c.y[i]),
                        zf = static_cast<float>(c.z[i]);
            if (fl & ASTRO_FLAG_OCCUPIED)
                atomicAdd(&tcount[grid_nearest(T, xf, yf, zf)], 1u);
        });
    }
};

template <class P>
struct KThermal {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> c;
        Counters* ctr;
        Grid T;
        double relax;   // -expm1(-g dt_sub / cap), the same for every cell of the substep
        double cap;
        Rec<P>* rec;
        const int32_t* spos;
        int update_rec;
        const uint32_t* tcount;   // DIRECT_FIRST: occupied entries per voxel
        uint8_t* talone;          // DIRECT: per occupied physical entry, 1 alone, 2 shared
        int mode;
        int parity;               // the substep's parity, for Counters::tqueued
        int last;                 // the tick's last substep
        Grid irr;                 // deposit non-null: also stamp occluders (KStamp) with the final flags
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [c, ctr, T, relax, cap, rec, spos, update_rec, tcount, talone, mode, parity, last, irr] = a;

    // This substep reports in tqueued[parity] whether any deposit waits in
    // the accumulator, and clears the other parity's report, which the
    // previous substep's apply pass has read.
    if (vg.block == 0 && threadIdx.x == 0) ctr->tqueued[parity ^ 1] = 0u;
    const int n = ctr->nphys;
    // The entry's flags after this substep (0 while unoccupied) and its
    // position, for the occluder stamp.
    uint32_t fin = 0u;
    float xf = 0.0f, yf = 0.0f, zf = 0.0f;
    // Returns whether the cell left a deposit in the accumulator.
    const auto body = [&](const int i) -> bool {
    // Flags and position are loaded together, ahead of the branch on the
    // flags, so one memory latency precedes the voxel read; a cell already
    // awake and alive fetches its energy beside that read (a cell that
    // ignites below fetches it then).
    const uint32_t fl = c.flags[i];
    xf = static_cast<float>(c.x[i]);
    yf = static_cast<float>(c.y[i]);
    zf = static_cast<float>(c.z[i]);
    if (!(fl & ASTRO_FLAG_OCCUPIED)) return false;
    fin = fl;
    const uint32_t idx = grid_nearest(T, xf, yf, zf);
    const bool early = (fl & ASTRO_FLAG_AWAKE) && (fl & ASTRO_FLAG_ALIVE);
    double e0 = early ? c.energy[i] : 0.0;
    const float t_grid = T.value[idx];
    uint8_t share = 2;
    if (mode == THERMAL_DIRECT_FIRST) {
        share = tcount[idx] == 1u ? 1 : 2;
        talone[i] = share;
    } else if (mode == THERMAL_DIRECT) {
        share = talone[i];
    }
    bool awake = (fl & ASTRO_FLAG_AWAKE) != 0u;
    uint32_t nf = fl;
    if (!awake && static_cast<double>(t_grid) >= ASTRO_CELL_TEMP_SETPOINT) {
        awake = true;
        nf = fl | ASTRO_FLAG_AWAKE;
        c.flags[i] = nf;
        fin = nf;
    }
    const double t_cell = awake ? ASTRO_CELL_TEMP_SETPOINT : static_cast<double>(t_grid);
    // nothing reads temp_cell between substeps: the last one's value stands
    if (last) c.temp_cell[i] = static_cast<float>(t_cell);
    if (!awake || !(fl & ASTRO_FLAG_ALIVE)) return false;
    if (!early) e0 = c.energy[i];
    // lumped_exchange_energy (src/sim/thermal.cuh:80-86) with its relaxation factor
    // taken once per substep on the host.
    const double e = cap > 0.0 ? cap * (t_cell - static_cast<double>(t_grid)) * relax : 0.0;
    double energy = e0 - e;
    double moved = e;
    if (energy < 0.0) {
        moved = e0;
        energy = 0.0;
    } else if (energy > ASTRO_CELL_ENERGY_MAX) {
        moved = e0 - ASTRO_CELL_ENERGY_MAX;
        energy = ASTRO_CELL_ENERGY_MAX;
    }
    c.energy[i] = energy;
    bool queued = false;
    if (share == 1) {
        const unsigned long long raw = static_cast<unsigned long long>(to_fixed(moved / cap, T.scale));
        if (raw != 0ull) T.value[idx] = t_grid + static_cast<float>(from_fixed(raw, T.scale));
    } else {
        deposit_add(&T.deposit[idx], moved / cap, T.scale);
        queued = true;
    }
    if (energy <= 0.0) {
        c.flags[i] = nf & ~ASTRO_FLAG_ALIVE;
        fin = nf & ~ASTRO_FLAG_ALIVE;
        c.death_cause[i] = ASTRO_DEATH_STARVED;
        c.emit_power[i] = 0.0f;
        if (update_rec) rec[spos[i]].meta &= ~REC_ALIVE;
    }
    return queued;
    };
    xcd_tiles(vg, n, [&](const int i, const bool valid) {
        fin = 0u;
        const bool q = valid && body(i);
        if (__ballot(q) != 0ull && (threadIdx.x & 63) == 0 && ctr->tqueued[parity] == 0u) ctr->tqueued[parity] = 1u;
        // occlusion_stamp_kernel (src/sim/emission.cu:17-27) on the flags the
        // thermal step leaves, as the emission stage that follows reads them
        if (irr.deposit && (fin & ASTRO_FLAG_OCCUPIED)) {
            const double area = (fin & ASTRO_FLAG_ALIVE) ? ASTRO_CELL_CROSS_SECTION
                                                         : ASTRO_CELL_CROSS_SECTION * (1.0 - ASTRO_CELL_ALBEDO_DEAD);
            deposit_add(&irr.deposit[grid_nearest(irr, xf, yf, zf)], area, irr.scale);
        }
    });
    }
};

// Folds a field's fixed-point accumulator into its values at the cells that
// may have deposited (flush_kernel, src/fields/diffuse.cu:51-60, made sparse).
// Every depositor exchanges its voxel's accumulator with zero; exactly one of
// them receives the total and applies it, so the voxel sees the same single
// `value += float(from_fixed(total))` as a dense flush. AWAKE marks every
// thermostat depositor (thermal.cu:130), dead or alive afterwards. On the
// DIRECT path only cells of shared voxels take part, and the first pass of the
// tick also clears the voxel counts. A substep whose thermal pass queued no
// deposit has nothing to fold, and the pass returns at once.
template <class P>
struct KApplyAwake {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> c;
        const Counters* ctr;
        Grid g;
        uint32_t* tcount;
        const uint8_t* talone;
        int mode;
        int parity;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [c, ctr, g, tcount, talone, mode, parity] = a;

    if (mode != THERMAL_DIRECT_FIRST && ctr->tqueued[parity] == 0u) return;   // uniform over the grid

    const int n = ctr->nphys;
    const auto body = [&](const int i) {
    // Flags and position are loaded together (in bounds for i below nphys).
    const uint32_t fl = c.flags[i];
    const float xf = static_cast<float>(c.x[i]), yf = static_cast<float>(c.y[i]), zf = static_cast<float>(c.z[i]);
    const uint8_t share = mode == THERMAL_ATOMIC ? 2 : talone[i];
    if (!(fl & ASTRO_FLAG_OCCUPIED)) return;
    if (mode == THERMAL_DIRECT_FIRST) tcount[grid_nearest(g, xf, yf, zf)] = 0u;
    if (share != 2 || !(fl & ASTRO_FLAG_AWAKE)) return;
    const uint32_t idx = grid_nearest(g, xf, yf, zf);
    const unsigned long long raw = atomicExch(&g.deposit[idx], 0ull);
    if (raw != 0ull) g.value[idx] += static_cast<float>(from_fixed(raw, g.scale));
    };
    xcd_tiles(vg, n, [&](const int i, const bool valid) { if (valid) body(i); });
    }
};

// ---------------------------------------------------------------------------
// Stand-alone feeding / taxis stage (emission absorption, taxis_step) over
// physical entries; the tick runs the same body inside the fused force pass.
// ---------------------------------------------------------------------------
template <class P, bool FEED, bool TAXIS>
struct KFeedTaxis {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> c;
        const Counters* ctr;
        const int32_t* spos;
        const double* dirs;
        float ambient;
        double dt;
        double max_power;
        uint64_t rng_seed;
        int philox;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [c, ctr, spos, dirs, ambient, dt, max_power, rng_seed, philox] = a;

    const int n = ctr->nphys;
    const RngCtx rng{rng_seed, &ctr->rng_tick, philox};
    xcd_tiles(vg, n, [&](const int i, const bool valid) {
        if (!valid) return;
        const uint32_t f = c.flags[i];
        if (!(f & ASTRO_FLAG_OCCUPIED) || !(f & ASTRO_FLAG_ALIVE)) return;
        const double directional = FEED ? dirs[spos[i]] : 0.0;
        feed_taxis_cell<P, FEED, TAXIS>(c, i, f, directional, ambient, dt, max_power, rng);
    });
    }
};

// ---------------------------------------------------------------------------
// flash_kernel (src/sim/flash.cu:31-56): every awake charged cell discharges
// at the flash power toward the objective. Its recoil drives the ship, not
// the cell (emit_power 0); the photon momentum and the energy go to the
// fixed-point audit.
// ---------------------------------------------------------------------------
template <class P>
struct KFlash {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> c;
        const Counters* ctr;
        double dt;
        double power;   // World::petrova_flash_power [W]
        unsigned long long* accum;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [c, ctr, dt, power, accum] = a;
        xcd_tiles(vg, ctr->nphys, [&](const int i, const bool valid) {
            if (!valid) return;
            const uint32_t f = c.flags[i];
            if (!(f & ASTRO_FLAG_OCCUPIED) || !(f & ASTRO_FLAG_ALIVE) || !(f & ASTRO_FLAG_AWAKE)) return;
            const double e = c.energy[i];
            c.emit_power[i] = 0.0f;
            if (e <= 0.0) return;
            double de = power * dt;
            if (de > e) de = e;
            c.energy[i] = e - de;
            deposit_add(&accum[2], -(de / ASTRO_C_LIGHT), ASTRO_FLASH_IMPULSE_SCALE);
            deposit_add(&accum[3], de, ASTRO_FLASH_ENERGY_SCALE);
        });
    }
};

// ---------------------------------------------------------------------------
// field_sample_kernel (src/sim/integrator.cu:20-30) and the body-force half of
// forces_kernel (:40-58), per physical entry; the force goes to the cell's
// sorted position for the contact walk to add to. With FEED / TAXIS the same
// pass first runs feed_kernel / taxis_kernel (src/sim/emission.cu:81-115,
// src/sim/taxis.cu:13-87) on the cell, which read no other cell.
// ---------------------------------------------------------------------------
template <class P, bool FEED, bool TAXIS>
struct KSampleBase {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> c;
        const Counters* ctr;
        Grid T;
        Grid C;
        Grid N;
        int c_on;
        int n_on;
        int gaxis;
        const int32_t* spos;
        Force<P>* fs;
        const double* dirs;
        float ambient;
        double dt;
        double max_power;
        uint64_t rng_seed;
        int philox;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [c, ctr, T, C, N, c_on, n_on, gaxis, spos, fs, dirs, ambient, dt, max_power, rng_seed,
                     philox] = a;

    const int n = ctr->nphys;
    const RngCtx rng{rng_seed, &ctr->rng_tick, philox};
    xcd_tiles(vg, n, [&](const int i, const bool valid) {
        if (!valid) return;
        // Slot, flags and position are loaded together (in bounds for i
        // below nphys); nothing before the force sample moves the cell.
        const uint32_t slot = c.lslot[i];
        uint32_t f = c.flags[i];
        const float xf = static_cast<float>(c.x[i]), yf = static_cast<float>(c.y[i]),
                    zf = static_cast<float>(c.z[i]);
        if (slot == NO_SLOT) return;
        if constexpr (FEED || TAXIS) {
            // feeding and taxis first, as their own pass would (no other cell
            // is read), then the force from the updated state
            if ((f & ASTRO_FLAG_OCCUPIED) && (f & ASTRO_FLAG_ALIVE)) {
                const double directional = FEED ? dirs[spos[i]] : 0.0;
                feed_taxis_cell<P, FEED, TAXIS>(c, i, f, directional, ambient, dt, max_power, rng);
                f = c.flags[i];
            }
        }
        const Force<double> F = sample_force_cell<P>(c, i, f, xf, yf, zf, T, C, N, c_on, n_on, gaxis);
        fs[spos[i]] = Force<P>{static_cast<P>(F.x), static_cast<P>(F.y), static_cast<P>(F.z)};
    });
    }
};

// ---------------------------------------------------------------------------
// integrate_kernel (src/sim/integrator.cu:79-150).
// ---------------------------------------------------------------------------
// Books one live cell's CO2 demand against its nearest voxel: demand_kernel
// (src/sim/lifecycle.cu:41-52) for one cell. Records the voxel (NO_VOXEL if
// none) and the CO2 it saw there, for the uptake.
//   The reference adds each cell's fixed-point demand into a 64-bit total per
// voxel. Every cell of one voxel reads the same concentration from the same
// field, so every one of them asks the same fixed-point amount d, and the
// voxel's total is its count of booking cells times d (modulo 2^64, as the
// reference's wrapping adds). Here a cell only counts itself into a 32-bit
// array -- half the width, so twice the voxels on each cache line, in the
// booking, in the uptake's read and in the clear -- and the uptake forms
// count * d from the amount it asked itself.
constexpr uint32_t NO_VOXEL = 0xFFFFFFFFu;
// Set on the one cell per shared voxel whose uptake deposit found the
// accumulator empty: that cell alone flushes the voxel. Voxel indices stay
// below 2^31 - 1.
constexpr uint32_t VOX_OWNER = 0x80000000u;

struct DemandArgs {
    Grid C;
    double dt_bio, gv;
    uint32_t* demand;   // booking cells per voxel
    uint32_t* vox;
    float* avail;
};

__device__ __forceinline__ void book_demand(const DemandArgs& a, int i, bool live, float x, float y, float z) {
    uint32_t v = NO_VOXEL;
    if (live) {
        const uint32_t idx = grid_nearest(a.C, x, y, z);
        const float cv = a.C.value[idx];
        const double want = co2_uptake_rate(static_cast<double>(cv)) * a.dt_bio;
        if (want > 0.0) {
            const int64_t fxd = to_fixed(want / a.gv, a.C.scale);
            if (fxd != 0) atomicAdd(&a.demand[idx], 1u);
            v = idx;
            a.avail[i] = cv;
        }
    }
    a.vox[i] = v;
}

// integrate_kernel (src/sim/integrator.cu:79-150). With DEMAND the same pass
// then books the moved cell's CO2 demand (lifecycle's first kernel, which
// reads nothing the steps in between change once the CO2 field has diffused).
template <class P, bool DEMAND, bool PHILOX>
struct KIntegrate {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> c;
        const Counters* ctr;
        const int32_t* spos;
        const Force<P>* fs;
        TickParams t;
        DemandArgs dm;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [c, ctr, spos, fs, t, dm] = a;

    const int n = ctr->nphys;
    const auto body = [&](const int i) {
    // The entry's own arrays are all loaded with its flags (in bounds for i
    // below nphys); only the force, found through the sorted position, waits
    // for the occupancy test.
    const uint32_t fl = c.flags[i];
    const int32_t sp = spos[i];
    const float t_local_f = c.t_local[i];
    const double biomass = c.biomass[i], energy = c.energy[i];
    double px = static_cast<double>(c.x[i]), py = static_cast<double>(c.y[i]), pz = static_cast<double>(c.z[i]);
    double vx = static_cast<double>(c.vx[i]), vy = static_cast<double>(c.vy[i]), vz = static_cast<double>(c.vz[i]);
    const uint64_t id = c.id[i];
    // A counter stream needs no per-cell state (Philox: seed, tick, id).
    const uint64_t rng_state = PHILOX ? 0ull : c.rng_state[i];
    if (!(fl & ASTRO_FLAG_OCCUPIED)) {
        if constexpr (DEMAND) dm.vox[i] = NO_VOXEL;
        return;
    }
    const double dt = t.dt;
    const double t_local = static_cast<double>(t_local_f);
    const double mass = cell_mass(biomass, energy);
    const bool awake = (fl & ASTRO_FLAG_AWAKE) != 0u;
    const double t_drag = awake ? ASTRO_CELL_TEMP_SETPOINT : t_local;
    const Force<P> F = fs[sp];
    auto r = [&] {
        if constexpr (PHILOX) return Philox(t.rng_seed, RNG_STAGE_INTEGRATE, ctr->rng_tick, id);
        else return cell_rng(rng_state, id);
    }();
    // The three pairs are drawn in the reference's order (x, y, z); each is
    // used as soon as it is drawn, which keeps fewer values live.
    const bool noise = t.thermal_noise != 0;
    if constexpr (std::is_same_v<P, float>) {
        // FP32 storage: float drag, coefficients and Gaussians (see OuF)
        float gf = drag_coefficient_f(static_cast<float>(t_drag));
        if (fl & ASTRO_FLAG_STUCK) gf *= static_cast<float>(ASTRO_WALL_STUCK_DRAG_MULT);
        const float dtf = static_cast<float>(dt);
        const OuF oc = ou_coefficients_f(static_cast<float>(mass), gf, static_cast<float>(t_drag), dtf);
        float z1, z2;
        gaussian_pair_f(r, z1, z2);
        ou_advance_axis_f(px, vx, static_cast<float>(F.x), gf, oc, dtf, noise, z1, z2);
        gaussian_pair_f(r, z1, z2);
        ou_advance_axis_f(py, vy, static_cast<float>(F.y), gf, oc, dtf, noise, z1, z2);
        gaussian_pair_f(r, z1, z2);
        ou_advance_axis_f(pz, vz, static_cast<float>(F.z), gf, oc, dtf, noise, z1, z2);
    } else {
        double gamma = drag_coefficient(t_drag);
        if (fl & ASTRO_FLAG_STUCK) gamma *= ASTRO_WALL_STUCK_DRAG_MULT;
        const Ou oc = ou_coefficients(mass, gamma, t_drag, dt);
        double z1, z2;
        r.gaussian_pair(z1, z2);
        ou_advance_axis(px, vx, F.x, gamma, oc, dt, noise, z1, z2);
        r.gaussian_pair(z1, z2);
        ou_advance_axis(py, vy, F.y, gamma, oc, dt, noise, z1, z2);
        r.gaussian_pair(z1, z2);
        ou_advance_axis(pz, vz, F.z, gamma, oc, dt, noise, z1, z2);
    }
    const double a = ASTRO_CELL_RADIUS;
    bool kept = apply_boundary_axis(px, vx, 0.5 * t.chamber_w, a, t.boundary_x);
    kept = apply_boundary_axis(py, vy, 0.5 * t.chamber_h, a, t.boundary_y) && kept;
    const double z_before = pz;
    apply_boundary_axis(pz, vz, 0.5 * t.chamber_d, a, ASTRO_BOUNDARY_REFLECTING);
    const bool touched = (pz != z_before);
    uint32_t nf = fl;
    if (t.adhesion) {
        if (touched && !(fl & ASTRO_FLAG_STUCK)) {
            // The reference re-derives the stream from the state it just
            // stored (integrator.cu:124-127); r.state is that state. A
            // counter stream takes its next draws.
            const bool sticks = r.uniform01d() < ASTRO_WALL_STICKINESS;
            if (sticks) nf |= ASTRO_FLAG_STUCK;
        } else if (fl & ASTRO_FLAG_STUCK) {
            const double away = (pz < 0.0) ? F.z : -F.z;
            if (away > ASTRO_WALL_RELEASE_FORCE) nf &= ~ASTRO_FLAG_STUCK;
        }
    }
    if constexpr (!PHILOX) c.rng_state[i] = r.state;
    if (!kept) {
        nf &= ~ASTRO_FLAG_ALIVE;
        c.death_cause[i] = ASTRO_DEATH_CULLED;
        c.emit_power[i] = 0.0f;
    }
    c.flags[i] = nf;
    const P sx = static_cast<P>(px), sy = static_cast<P>(py), sz = static_cast<P>(pz);
    c.x[i] = sx;
    c.y[i] = sy;
    c.z[i] = sz;
    c.vx[i] = static_cast<P>(vx);
    c.vy[i] = static_cast<P>(vy);
    c.vz[i] = static_cast<P>(vz);
    c.age_s[i] += static_cast<float>(dt);
    if constexpr (DEMAND)
        book_demand(dm, i, (nf & ASTRO_FLAG_ALIVE) != 0u, static_cast<float>(sx), static_cast<float>(sy),
                    static_cast<float>(sz));
    };
    xcd_tiles(vg, n, [&](const int i, const bool valid) { if (valid) body(i); });
    }
};

// ---------------------------------------------------------------------------
// lifecycle_step (src/sim/lifecycle.cu:208-294), host syncs replaced by
// device counters.
// ---------------------------------------------------------------------------
// demand_kernel (lifecycle.cu:41-52) on its own, for the stage entry point.
template <class P>
struct KDemand {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> c;
        const Counters* ctr;
        DemandArgs dm;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [c, ctr, dm] = a;

    const int n = ctr->nphys;
    xcd_tiles(vg, n, [&](const int i, const bool valid) {
        if (!valid) return;
        const uint32_t f = c.flags[i];
        book_demand(dm, i, (f & ASTRO_FLAG_OCCUPIED) && (f & ASTRO_FLAG_ALIVE), static_cast<float>(c.x[i]),
                    static_cast<float>(c.y[i]), static_cast<float>(c.z[i]));
    });
    }
};

// uptake_kernel (lifecycle.cu:62-86), death_kernel (:100-112) and
// mark_kernel (:119-135) in one pass; each touches only the cell's own state
// except the uptake's voxel.
//   * Uptake. Nothing in this pass reads the CO2 field itself (the demand
//     pass stored the concentration each cell saw), so the flush that follows
//     every uptake in the reference may be done per voxel as soon as that
//     voxel's total is known. A cell whose own demand is its voxel's whole
//     demand is the voxel's only depositor (demands are nonnegative and a
//     zero-demand cell's take rounds to a zero deposit), so it applies
//     `value += float(from_fixed(deposit))` itself, with no accumulator. In a
//     shared voxel the deposits accumulate atomically; they are all negative,
//     so exactly one depositor sees the accumulator still at zero, and it is
//     marked the voxel's owner for the flush in the apply pass.
//   * Divide and drop flags are set by reference slot, each with its scan-tile
//     total.
template <class P>
struct KUptakeMark {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> c;
        Counters* ctr;
        DemandArgs dm;
        int c_on;
        int disposition;
        double quota;
        int32_t* dflag;
        int32_t* dsum;
        int32_t* drop;
        int32_t* xsum;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [c, ctr, dm, c_on, disposition, quota, dflag, dsum, drop, xsum] = a;

    const int n = ctr->nphys;
    xcd_tiles(vg, n, [&](const int i, const bool valid) {
        bool divides = false, dead = false;
        if (valid) {
            // Every per-entry value the pass reads is loaded up front (all in
            // bounds for i below nphys), so only the voxel's demand waits on
            // a second memory latency.
            const uint32_t idx = c_on ? dm.vox[i] : NO_VOXEL;
            const float avail_f = c_on ? dm.avail[i] : 0.0f;
            uint32_t f = c.flags[i];
            const float t_local_f = c.t_local[i];
            double held = c.co2_held[i];
            const uint32_t s = c.lslot[i];
            if (c_on) {
                if (idx != NO_VOXEL) {
                    const double available = static_cast<double>(avail_f);
                    const double want = co2_uptake_rate(available) * dm.dt_bio;
                    // the voxel's total: its booking cells, each asking what
                    // this cell asked (see book_demand)
                    const int64_t own = to_fixed(want / dm.gv, dm.C.scale);
                    const unsigned long long raw =
                        static_cast<unsigned long long>(dm.demand[idx]) * static_cast<unsigned long long>(own);
                    const double asked = from_fixed(raw, dm.C.scale);
                    double taken = want;
                    if (asked > available && asked > 0.0) taken = want * (available / asked);
                    if (taken > 0.0) {
                        held += taken;
                        c.co2_held[i] = held;
                        const int64_t fxd = to_fixed(-taken / dm.gv, dm.C.scale);
                        if (fxd != 0) {
                            if (static_cast<int64_t>(raw) == own) {
                                dm.C.value[idx] += static_cast<float>(
                                    from_fixed(static_cast<unsigned long long>(fxd), dm.C.scale));
                            } else {
                                const unsigned long long old =
                                    atomicAdd(&dm.C.deposit[idx], static_cast<unsigned long long>(fxd));
                                if (old == 0ull) dm.vox[i] = idx | VOX_OWNER;
                            }
                        }
                    }
                }
            }
            if ((f & ASTRO_FLAG_OCCUPIED) && (f & ASTRO_FLAG_ALIVE) &&
                static_cast<double>(t_local_f) > ASTRO_CELL_LETHAL_TEMP) {
                f &= ~ASTRO_FLAG_ALIVE;
                c.flags[i] = f;
                c.death_cause[i] = ASTRO_DEATH_OVERHEATED;
                c.emit_power[i] = 0.0f;
                c.energy[i] = disposition == ASTRO_STORE_RETAIN ? c.energy[i] : 0.0;
            }
            const bool occ = (f & ASTRO_FLAG_OCCUPIED) != 0u, alive = (f & ASTRO_FLAG_ALIVE) != 0u;
            divides = occ && alive && held >= quota;
            dead = occ && !alive;
            if (s != NO_SLOT) {
                // Both arrays and their tile totals start the tick cleared.
                if (divides) {
                    dflag[s] = 1;
                    atomicAdd(&dsum[s >> SCAN_TILE_SHIFT], 1);
                }
                if (!(occ && alive)) {   // compaction keeps occupied && alive
                    drop[s] = 1;
                    atomicAdd(&xsum[s >> SCAN_TILE_SHIFT], 1);
                }
            }
        }
        wave_count_add(&ctr->births, divides);
        wave_count_add(&ctr->deads, dead);
    });
    }
};

// Whether this tick compacts: lifecycle.cu:289 (compaction on and at least
// one occupied dead slot), or a direct cell_store_compact request.
__device__ __forceinline__ bool compacting_now(const Counters* ctr, int compaction, int forced) {
    return forced != 0 || (compaction != 0 && ctr->deads > 0);
}

// Both lifecycle scans in one launch, over reference slots [0, count): blocks
// [0, T) scan the divide flags (when anything divides), blocks [T, 2T) the
// drop flags (when compacting), T = tiles(capacity). The drop scan also lists
// the dropped slots in ascending order for the renumbering pass.
struct KLifeScan {
    struct Args {
        Counters* ctr;
        int tiles;
        int compaction;
        int forced;
        const int32_t* dflag;
        const int32_t* dsum;
        int32_t* prefix;
        const int32_t* drop;
        const int32_t* xsum;
        int32_t* dprefix;
        uint32_t* dlist;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [ctr, tiles, compaction, forced, dflag, dsum, prefix, drop, xsum, dprefix, dlist] = a;

    const int n = ctr->count;
    if (vg.block < tiles) {
        if (ctr->births <= 0) return;
        scan_tile(
            n, vg.block, [&](int i) { return load4_i32(dflag, i, n); },
            [&](int i, int4 ex, int4) { store4_i32(prefix, i, n, ex); }, dsum, &ctr->div_total, &ctr->error_flags);
    } else {
        if (!compacting_now(ctr, compaction, forced)) return;
        scan_tile(
            n, vg.block - tiles, [&](int i) { return load4_i32(drop, i, n); },
            [&](int i, int4 ex, int4 v) {
                store4_i32(dprefix, i, n, ex);
                if (v.x) dlist[ex.x] = static_cast<uint32_t>(i);
                if (v.y) dlist[ex.y] = static_cast<uint32_t>(i + 1);
                if (v.z) dlist[ex.z] = static_cast<uint32_t>(i + 2);
                if (v.w) dlist[ex.w] = static_cast<uint32_t>(i + 3);
            },
            xsum, &ctr->drop_total, &ctr->error_flags);
    }
    }
};

// Dropped slots a block keeps in LDS for the renumbering binary search; with
// more drops than this the pass reads the dense prefix instead.
constexpr int DROP_LDS_MAX = 4096;

// divide_kernel (lifecycle.cu:140-204) and the renumbering half of
// cell_store_compact (cell_store.cu:414-456) in one pass over physical
// entries.
//   * A daughter's reference slot is count + prefix (the prefix over
//     reference-slot order) and its id next_id + prefix; physically it is
//     appended after the last entry. Compaction then moves it down by the
//     number of dropped slots, all of which lie below count.
//   * A survivor's slot moves down by the number of dropped slots below it; a
//     dropped entry becomes a tombstone. Nothing is copied: storage order is
//     the spatial order, and the next re-sort packs tombstones out.
template <class P>
struct KLifeApply {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> c;
        Counters* ctr;
        const int32_t* prefix;
        const int32_t* drop;
        const int32_t* dprefix;
        const uint32_t* dlist;
        TickParams t;
        int forced;
        int c_on;
        DemandArgs dm;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [c, ctr, prefix, drop, dprefix, dlist, t, forced, c_on, dm] = a;

    __shared__ uint32_t s_drop[DROP_LDS_MAX];
    const int np = ctr->nphys;
    const int n = ctr->count;
    const int births = forced ? 0 : ctr->births;
    const bool compacting = compacting_now(ctr, t.compaction, forced);
    const int ndrop = compacting ? ctr->drop_total : 0;
    const bool in_lds = ndrop <= DROP_LDS_MAX;
    if (compacting && in_lds) {
        for (int k = threadIdx.x; k < ndrop; k += BLOCK) s_drop[k] = dlist[k];
        __syncthreads();
    }
    if (births > 0 || compacting || c_on) {
        xcd_tiles(vg, np, [&](const int i, const bool valid) {
            if (!valid) return;
            // The voxel record, the slot and (when anything divides) the
            // division test's flags and CO2 are loaded together (in bounds
            // for i below nphys).
            const uint32_t v = c_on ? dm.vox[i] : NO_VOXEL;
            const uint32_t s = c.lslot[i];
            const uint32_t f = births > 0 ? c.flags[i] : 0u;
            const double held = births > 0 ? c.co2_held[i] : 0.0;
            if (c_on) {
                // The CO2 flush of shared voxels (lifecycle.cu:235) by their
                // owners, and the demand cleared for the next tick.
                if (v != NO_VOXEL) {
                    const uint32_t idx = v & ~VOX_OWNER;
                    if (v & VOX_OWNER) {
                        const unsigned long long raw = dm.C.deposit[idx];
                        if (raw != 0ull) {
                            dm.C.value[idx] += static_cast<float>(from_fixed(raw, dm.C.scale));
                            dm.C.deposit[idx] = 0ull;
                        }
                    }
                    dm.demand[idx] = 0u;
                }
            }
            if (s == NO_SLOT) return;
            uint32_t below = 0u;
            if (compacting) {
                bool dropped;
                if (in_lds) {
                    int lo = 0, hi = ndrop;   // first entry >= s
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (s_drop[mid] < s) lo = mid + 1;
                        else hi = mid;
                    }
                    dropped = lo < ndrop && s_drop[lo] == s;
                    below = static_cast<uint32_t>(lo);
                } else {
                    dropped = drop[s] != 0;
                    below = static_cast<uint32_t>(dprefix[s]);
                }
                if (dropped) {
                    c.lslot[i] = NO_SLOT;
                    c.flags[i] = 0u;
                    return;
                }
                if (below) c.lslot[i] = s - below;
            }
            if (births <= 0) return;
            if (!(f & ASTRO_FLAG_OCCUPIED) || !(f & ASTRO_FLAG_ALIVE)) return;
            if (!(held >= t.quota)) return;
            const int32_t pre = prefix[s];
            const int32_t lslot = n + pre;
            if (lslot >= ctr->capacity) return;
            const int32_t slot = np + pre;
            c.lslot[slot] = static_cast<uint32_t>(lslot - ndrop);
            const uint64_t did = ctr->next_id + static_cast<uint64_t>(pre);
            const double half0 = 0.5 * (c.energy[i] - ASTRO_LIFE_DIVISION_ENERGY_COST);
            const double half = half0 > 0.0 ? half0 : 0.0;
            c.energy[i] = half;
            c.energy[slot] = half;
            c.co2_held[i] = 0.0;
            c.co2_held[slot] = 0.0;
            c.biomass[slot] = ASTRO_CELL_MASS_DRY;
            // division_axis (src/sim/lifecycle.cuh:87-99)
            const uint64_t h = splitmix64(splitmix64(did));
            const double u = static_cast<double>(h >> 33) * (1.0 / 2147483648.0) - 1.0;
            const double phi = static_cast<double>(h & 0xFFFFFFFFull) * (ASTRO_TWO_PI / 4294967296.0);
            const double sq = sqrt(u * u < 1.0 ? 1.0 - u * u : 0.0);
            const double ax = sq * cos(phi), ay = sq * sin(phi), az = u;
            const double a = t.division_offset;
            const double px = static_cast<double>(c.x[i]), py = static_cast<double>(c.y[i]),
                         pz = static_cast<double>(c.z[i]);
            double dx = px + a * ax, dy = py + a * ay, dz = pz + a * az;
            double qx = px - a * ax, qy = py - a * ay, qz = pz - a * az;
            const double hx = 0.5 * t.chamber_w - ASTRO_CELL_RADIUS;
            const double hy = 0.5 * t.chamber_h - ASTRO_CELL_RADIUS;
            const double hz = 0.5 * t.chamber_d - ASTRO_CELL_RADIUS;
            auto cl = [](double v, double hh) { return v < -hh ? -hh : (v > hh ? hh : v); };
            // In FP32 storage the reference's read-back of the stored daughter
            // coordinate before clamping is a rounding of the same value.
            dx = cl(static_cast<double>(static_cast<P>(dx)), hx);
            dy = cl(static_cast<double>(static_cast<P>(dy)), hy);
            dz = cl(static_cast<double>(static_cast<P>(dz)), hz);
            qx = cl(static_cast<double>(static_cast<P>(qx)), hx);
            qy = cl(static_cast<double>(static_cast<P>(qy)), hy);
            qz = cl(static_cast<double>(static_cast<P>(qz)), hz);
            c.x[slot] = static_cast<P>(dx); c.y[slot] = static_cast<P>(dy); c.z[slot] = static_cast<P>(dz);
            c.x[i] = static_cast<P>(qx); c.y[i] = static_cast<P>(qy); c.z[i] = static_cast<P>(qz);
            c.vx[slot] = c.vx[i]; c.vy[slot] = c.vy[i]; c.vz[slot] = c.vz[i];
            c.id[slot] = did;
            c.flags[slot] = f;
            c.death_cause[slot] = ASTRO_DEATH_NONE;
            c.temp_cell[slot] = c.temp_cell[i];
            c.emit_power[slot] = 0.0f;
            c.dir_x[slot] = c.dir_x[i]; c.dir_y[slot] = c.dir_y[i]; c.dir_z[slot] = c.dir_z[i];
            c.age_s[slot] = 0.0f;
            // a counter stream is keyed by the daughter's own id: no state to split
            c.rng_state[slot] = t.philox ? 0ull : pcg_split(c.rng_state[i], did);
            c.taxis_memory[slot] = 0.0f;
            c.run_timer[slot] = 0.0f;
            c.t_local[slot] = c.t_local[i];
            c.irradiance[slot] = 0.0f;
            c.co2_local[slot] = c.co2_local[i];
            c.n2_local[slot] = c.n2_local[i];
        });
    }
    }
};

// The host bookkeeping of lifecycle_step and cell_store_compact
// (lifecycle.cu:251-294, cell_store.cu:449-455), on device, after the apply
// pass has read the counters.
struct KLifeFinalize {
    struct Args {
        Counters* ctr;
        int compaction;
        int forced;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [ctr, compaction, forced] = a;

    if (threadIdx.x != 0 || vg.block != 0) return;
    const int32_t n = ctr->count;
    const bool compacting = compacting_now(ctr, compaction, forced);
    const int32_t ndrop = compacting ? ctr->drop_total : 0;
    if (forced) {
        ctr->count = n - ndrop;
        ctr->live = ctr->count;
        return;
    }
    const int32_t deads = ctr->deads;
    const int32_t new_deaths = deads - ctr->dead_slots_prev;
    if (new_deaths > 0) ctr->deaths_total += new_deaths;
    int32_t b = ctr->births;
    if (b > 0) {
        if (n + b > ctr->capacity) b = ctr->capacity - n;
        if (b > 0) {
            ctr->divisions_total += b;
            ctr->next_id += static_cast<uint64_t>(b);
            ctr->nphys += b;
        }
    }
    if (b < 0) b = 0;
    ctr->count = n + b - ndrop;
    if (compacting) ctr->live = ctr->count;
    ctr->dead_slots_prev = compacting ? 0 : deads;
    ctr->births = 0;
    ctr->deads = 0;
    }
};

// A direct compaction request (cell_store_compact outside lifecycle): every
// live entry writes its own slot's drop flag and tile total afresh.
template <class P>
struct KCompactMark {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> c;
        const Counters* ctr;
        int32_t* drop;
        int32_t* xsum;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [c, ctr, drop, xsum] = a;

    const int np = ctr->nphys;
    xcd_tiles(vg, np, [&](const int i, const bool valid) {
        if (!valid) return;
        const uint32_t s = c.lslot[i];
        if (s == NO_SLOT) return;
        const uint32_t f = c.flags[i];
        const bool keep = (f & ASTRO_FLAG_OCCUPIED) && (f & ASTRO_FLAG_ALIVE);
        drop[s] = keep ? 0 : 1;
        if (!keep) atomicAdd(&xsum[s >> SCAN_TILE_SHIFT], 1);
    });
    }
};

// ---------------------------------------------------------------------------
// stats_cells_kernel (src/sim/stats.cu:27-43): per-element fixed point, summed
// in registers over the thread's cells, then one block reduction and one
// atomic per block and word (the sums are exact integers, so the order of
// the additions does not matter).
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

constexpr int STATS_WORDS = 6;

template <class P>
__global__ __launch_bounds__(BLOCK) void k_stats_cells(Cells<P> c, const Counters* ctr, unsigned long long* acc) {
    __shared__ unsigned long long s_part[BLOCK / 64][STATS_WORDS];
    const int n = ctr->nphys;
    unsigned long long v[STATS_WORDS] = {0, 0, 0, 0, 0, 0};
    xcd_tiles(VGrid{static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x)}, n, [&](const int i, const bool valid) {
        if (!valid) return;
        const uint32_t f = c.flags[i];
        if (!(f & ASTRO_FLAG_OCCUPIED)) return;
        const bool alive = (f & ASTRO_FLAG_ALIVE) != 0u, awake = (f & ASTRO_FLAG_AWAKE) != 0u;
        v[0] += static_cast<unsigned long long>(to_fixed(c.energy[i], ASTRO_DEPOSIT_SCALE_STATS));
        if (alive)
            v[1] += static_cast<unsigned long long>(
                to_fixed(static_cast<double>(c.temp_cell[i]), ASTRO_DEPOSIT_SCALE_STATS));
        v[2] += alive; v[3] += !alive; v[4] += awake; v[5] += !awake;
    });
#pragma unroll
    for (int k = 0; k < STATS_WORDS; ++k) v[k] = wave_sum_u64(v[k]);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < STATS_WORDS; ++k) s_part[threadIdx.x >> 6][k] = v[k];
    __syncthreads();
    if (threadIdx.x < STATS_WORDS) {
        unsigned long long t = 0;
        for (int w = 0; w < BLOCK / 64; ++w) t += s_part[w][threadIdx.x];
        int word = SW_ENERGY;
        switch (threadIdx.x) {
            case 1: word = SW_TEMP_CELL; break;
            case 2: word = SW_N_ALIVE; break;
            case 3: word = SW_N_DEAD; break;
            case 4: word = SW_N_AWAKE; break;
            case 5: word = SW_N_DORMANT; break;
            default: break;
        }
        if (t) atomicAdd(&acc[word], t);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define PROF_BEGIN(name) \
    if (L.prof) L.prof->begin(name)
#define PROF_END() \
    if (L.prof) L.prof->end()

template <class P>
void launch_spawn(const Launch& L, const WorldDev<P>& w, const SpawnArgs& a, int32_t base, int32_t lbase,
                  uint64_t first_id, uint64_t seed) {
    PROF_BEGIN("spawn");
    k_spawn<P><<<grid_for(a.count), BLOCK, 0, L.stream>>>(w.c, a, base, lbase, first_id, seed);
    PROF_END();
}

static_assert(GridStride<KThermal<double>> && GridStride<KLifeApply<float>>);
static_assert(!GridStride<KLifeScan> && !GridStride<KLifeFinalize>);

template <class P>
void launch_thermal_count(const Launch& L, Views<P> ws) {
    using K = KThermalCount<P>;
    Batch<K> b;
    for (const WorldDev<P>& w : ws) b.add({w.c, w.ctr, w.g[ASTRO_FIELD_TEMPERATURE], w.tcount}, cell_grid(w.pcap));
    launch_k<K>(L, "thermal_count", b.args, b.blocks);
}

template <class P>
void launch_thermal_substep(const Launch& L, Views<P> ws, Ticks ts, double dt_sub, bool update_rec, int mode,
                            int substep, bool last, bool stamp) {
    using K = KThermal<P>;
    Batch<K> b;
    for (size_t i = 0; i < ws.size(); ++i) {
        const WorldDev<P>& w = ws[i];
        const double cap = ts[i].T_cap;
        const double g = ASTRO_CONDUCTION_COEFF;
        const double relax = (cap <= 0.0 || g <= 0.0) ? 0.0 : -std::expm1(-g * dt_sub / cap);
        // chambers with a positive source stamp occluders (enqueue_light)
        const Grid irr = stamp && last && ts[i].source > 0.0f ? w.g[ASTRO_FIELD_IRRADIANCE] : Grid{};
        b.add({w.c, w.ctr, w.g[ASTRO_FIELD_TEMPERATURE], relax, cap, w.rec, w.spos, update_rec ? 1 : 0, w.tcount,
               w.talone, mode, substep & 1, last ? 1 : 0, irr},
              cell_grid(w.pcap));
    }
    launch_k<K>(L, "thermal", b.args, b.blocks);
}

template <class P>
void launch_apply_awake(const Launch& L, Views<P> ws, int field, int mode, int substep) {
    using K = KApplyAwake<P>;
    Batch<K> b;
    for (const WorldDev<P>& w : ws)
        b.add({w.c, w.ctr, w.g[field], w.tcount, w.talone, mode, substep & 1}, cell_grid(w.pcap));
    launch_k<K>(L, "deposit_apply", b.args, b.blocks);
}

template <class P, bool FEED, bool TAXIS>
void launch_feed_taxis_t(const Launch& L, Views<P> ws, Ticks ts, const char* name) {
    using K = KFeedTaxis<P, FEED, TAXIS>;
    Batch<K> b;
    for (size_t i = 0; i < ws.size(); ++i) {
        const WorldDev<P>& w = ws[i];
        b.add({w.c, w.ctr, w.spos, w.dirs, ts[i].ambient_irr, ts[i].dt, ts[i].petrova_max_power, ts[i].rng_seed,
               ts[i].philox},
              cell_grid(w.pcap));
    }
    launch_k<K>(L, name, b.args, b.blocks);
}

template <class P>
void launch_feed_taxis(const Launch& L, Views<P> ws, Ticks ts, bool feed, bool taxis) {
    if (feed && taxis) launch_feed_taxis_t<P, true, true>(L, ws, ts, "feed_taxis");
    else if (feed) launch_feed_taxis_t<P, true, false>(L, ws, ts, "feed");
    else if (taxis) launch_feed_taxis_t<P, false, true>(L, ws, ts, "taxis");
}

template <class P>
void launch_flash(const Launch& L, Views<P> ws, Ticks ts) {
    using K = KFlash<P>;
    Batch<K> b;
    for (size_t i = 0; i < ws.size(); ++i) b.add({ws[i].c, ws[i].ctr, ts[i].dt, ts[i].petrova_flash_power, ws[i].flash}, cell_grid(ws[i].pcap));
    launch_k<K>(L, "flash", b.args, b.blocks);
}

template <class P, bool FEED, bool TAXIS>
void launch_sample_t(const Launch& L, Views<P> ws, Ticks ts, const char* name) {
    using K = KSampleBase<P, FEED, TAXIS>;
    Batch<K> b;
    for (size_t i = 0; i < ws.size(); ++i) {
        const WorldDev<P>& w = ws[i];
        const TickParams& t = ts[i];
        b.add({w.c, w.ctr, w.g[ASTRO_FIELD_TEMPERATURE], w.g[ASTRO_FIELD_CO2], w.g[ASTRO_FIELD_N2],
               w.on[ASTRO_FIELD_CO2] ? 1 : 0, w.on[ASTRO_FIELD_N2] ? 1 : 0, t.gravity_axis, w.spos, w.fs, w.dirs,
               t.ambient_irr, t.dt, t.petrova_max_power, t.rng_seed, t.philox},
              cell_grid(w.pcap));
    }
    launch_k<K>(L, name, b.args, b.blocks);
}

template <class P>
void launch_cell_forces(const Launch& L, Views<P> ws, Ticks ts, bool feed, bool taxis) {
    if (feed && taxis) launch_sample_t<P, true, true>(L, ws, ts, "feed_taxis_sample");
    else if (feed) launch_sample_t<P, true, false>(L, ws, ts, "feed_sample");
    else if (taxis) launch_sample_t<P, false, true>(L, ws, ts, "taxis_sample");
    else launch_sample_t<P, false, false>(L, ws, ts, "sample_base");
    launch_contact_walk<P>(L, ws, ts);
}

template <class P>
DemandArgs demand_args(const WorldDev<P>& w, const TickParams& t) {
    return DemandArgs{w.g[ASTRO_FIELD_CO2], t.dt_bio, t.co2_gv, w.co2_demand, w.rin, w.co2_avail};
}

template <class P, bool DEMAND, bool PHILOX>
void launch_integrate_t(const Launch& L, Views<P> ws, Ticks ts, const char* name) {
    using K = KIntegrate<P, DEMAND, PHILOX>;
    Batch<K> b;
    for (size_t i = 0; i < ws.size(); ++i)
        b.add({ws[i].c, ws[i].ctr, ws[i].spos, ws[i].fs, ts[i], demand_args(ws[i], ts[i])}, cell_grid(ws[i].pcap));
    // FP64 layout: 168 VGPRs, 3 waves per SIMD; the FP32 layout's float
    // integrator fits 4
    launch_k<K, std::is_same_v<P, float> ? 4 : 3>(L, name, b.args, b.blocks);
}

template <class P>
void launch_integrate(const Launch& L, Views<P> ws, Ticks ts, bool demand) {
    // chambers of one launch agree on the generator (the ensemble checks it)
    if (ts[0].philox) {
        if (demand) launch_integrate_t<P, true, true>(L, ws, ts, "integrate_demand");
        else launch_integrate_t<P, false, true>(L, ws, ts, "integrate");
    } else {
        if (demand) launch_integrate_t<P, true, false>(L, ws, ts, "integrate_demand");
        else launch_integrate_t<P, false, false>(L, ws, ts, "integrate");
    }
}

// The division / compaction tail shared by the lifecycle and a direct
// compaction: both scans, then the apply pass, then the bookkeeping.
template <class P>
void launch_life_tail(const Launch& L, Views<P> ws, Ticks ts, bool forced) {
    Batch<KLifeScan> scan;
    Batch<KLifeApply<P>> apply;
    Batch<KLifeFinalize> fin;
    for (size_t i = 0; i < ws.size(); ++i) {
        const WorldDev<P>& w = ws[i];
        const TickParams& t = ts[i];
        const int tiles = scan_tiles_for(w.capacity);
        const int c_on = (!forced && w.on[ASTRO_FIELD_CO2]) ? 1 : 0;
        scan.add({w.ctr, tiles, t.compaction, forced ? 1 : 0, w.dflag, w.dsum, w.prefix, w.drop, w.xsum, w.dprefix,
                  w.dlist},
                 2 * tiles);
        apply.add({w.c, w.ctr, w.prefix, w.drop, w.dprefix, w.dlist, t, forced ? 1 : 0, c_on, demand_args(w, t)},
                  cell_grid(w.pcap));
        fin.add({w.ctr, t.compaction, forced ? 1 : 0}, 1);
    }
    launch_k<KLifeScan>(L, "life_scan", scan.args, scan.blocks);
    launch_k<KLifeApply<P>>(L, "life_apply", apply.args, apply.blocks);
    launch_k<KLifeFinalize>(L, "life_finalize", fin.args, fin.blocks);
}

template <class P>
void launch_lifecycle(const Launch& L, Views<P> ws, Ticks ts, bool demand_booked) {
    const bool c_on = ws[0].on[ASTRO_FIELD_CO2];
    if (c_on && !demand_booked) {
        using K = KDemand<P>;
        Batch<K> b;
        for (size_t i = 0; i < ws.size(); ++i) b.add({ws[i].c, ws[i].ctr, demand_args(ws[i], ts[i])}, cell_grid(ws[i].pcap));
        launch_k<K>(L, "co2_demand", b.args, b.blocks);
    }
    using K = KUptakeMark<P>;
    Batch<K> b;
    for (size_t i = 0; i < ws.size(); ++i) {
        const WorldDev<P>& w = ws[i];
        b.add({w.c, w.ctr, demand_args(w, ts[i]), c_on ? 1 : 0, ts[i].store_disposition, ts[i].quota, w.dflag, w.dsum,
               w.drop, w.xsum},
              tile_grid(w.capacity));
    }
    launch_k<K>(L, "uptake_mark", b.args, b.blocks);
    launch_life_tail<P>(L, ws, ts, false);
}

template <class P>
void launch_life_clear(const Launch& L, const WorldDev<P>& w) {
    const size_t tiles = static_cast<size_t>(scan_tiles_for(w.capacity));
    hipMemsetAsync(w.dflag, 0, sizeof(int32_t) * w.capacity, L.stream);
    hipMemsetAsync(w.drop, 0, sizeof(int32_t) * w.capacity, L.stream);
    hipMemsetAsync(w.dsum, 0, sizeof(int32_t) * tiles, L.stream);
    hipMemsetAsync(w.xsum, 0, sizeof(int32_t) * tiles, L.stream);
}

template <class P>
void launch_compact(const Launch& L, const WorldDev<P>& w) {
    launch_life_clear<P>(L, w);
    using K = KCompactMark<P>;
    launch_k<K>(L, "compact_mark", {{w.c, w.ctr, w.drop, w.xsum}}, {cell_grid(w.pcap)});
    TickParams t{};
    launch_life_tail<P>(L, Views<P>{&w, 1}, Ticks{&t, 1}, true);
}

template <class P>
void launch_stats_cells(const Launch& L, const WorldDev<P>& w) {
    const int g = std::min(cell_grid(w.pcap), 512);
    k_stats_cells<P><<<g, BLOCK, 0, L.stream>>>(w.c, w.ctr, w.stats);
}

#define INSTANTIATE(P)                                                                                          \
    template void launch_spawn<P>(const Launch&, const WorldDev<P>&, const SpawnArgs&, int32_t, int32_t,       \
                                  uint64_t, uint64_t);                                                          \
    template void launch_thermal_count<P>(const Launch&, Views<P>);                                             \
    template void launch_thermal_substep<P>(const Launch&, Views<P>, Ticks, double, bool, int, int, bool, bool); \
    template void launch_apply_awake<P>(const Launch&, Views<P>, int, int, int);                                \
    template void launch_feed_taxis<P>(const Launch&, Views<P>, Ticks, bool, bool);                             \
    template void launch_integrate<P>(const Launch&, Views<P>, Ticks, bool);                                    \
    template void launch_cell_forces<P>(const Launch&, Views<P>, Ticks, bool, bool);                            \
    template void launch_flash<P>(const Launch&, Views<P>, Ticks);                                                \
    template void launch_lifecycle<P>(const Launch&, Views<P>, Ticks, bool);                                    \
    template void launch_compact<P>(const Launch&, const WorldDev<P>&);                                          \
    template void launch_life_clear<P>(const Launch&, const WorldDev<P>&);                                      \
    template void launch_stats_cells<P>(const Launch&, const WorldDev<P>&);

INSTANTIATE(double)
INSTANTIATE(float)

}  // namespace astro_dev
