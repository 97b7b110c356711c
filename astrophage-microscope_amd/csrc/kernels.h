// kernels.h -- host-side launchers of the hot-path kernels.
//
// Each launcher enqueues its kernels on the world's stream and never
// synchronizes. Kernel grids are sized for the capacity; the live population
// count is read on the device, so a launch sequence can be captured once into
// a hipGraph and replayed.
#pragma once

#include <span>
#include <string>
#include <vector>

#include "astro_dev.h"
#include "dispatch.h"
#include "scan.h"

namespace astro_dev {

// Per-launch timing hook (HIP events on the world's stream).
struct Profiler {
    virtual void begin(const char* name) = 0;
    virtual void end() = 0;
    virtual ~Profiler() = default;
};

struct Launch {
    hipStream_t stream;
    Profiler*   prof;    // null unless profiling
    Recorder*   rec;     // null: launch now for one world; else record a batched launch
};

// Launches body K for each chamber: args[i] with blocks[i] blocks (a multiple
// of NUM_XCD for bodies that tile with xcd_tiles). Without a recorder there is
// one chamber and the launch happens now; with one, a single batched launch
// over all chambers is recorded.
template <class K>
concept GridStride = requires { K::stride; } && static_cast<bool>(K::stride);

// W > 0 caps the body's registers for W waves per SIMD.
template <class K, int W = 0>
void launch_k(const Launch& L, const char* name, const std::vector<typename K::Args>& args,
              std::vector<int> blocks) {
    if (!L.rec) {
        if (L.prof) L.prof->begin(name);
        if constexpr (W > 0) k_one_w<K, W><<<blocks[0], BLOCK, 0, L.stream>>>(args[0]);
        else k_one<K><<<blocks[0], BLOCK, 0, L.stream>>>(args[0]);
        if (L.prof) L.prof->end();
        return;
    }
    if constexpr (GridStride<K>) spread_blocks(blocks, L.rec->blocks_for(name));
    int total = 0;
    for (int b : blocks) total += b;
    if (total == 0) return;
    const size_t a_off = L.rec->put(args.data(), sizeof(typename K::Args) * args.size());
    const size_t m_off = L.rec->block_map(blocks);
    L.rec->launches.push_back({name, [=](hipStream_t s, const char* arena) {
        const auto* a = reinterpret_cast<const typename K::Args*>(arena + a_off);
        const int4* m = reinterpret_cast<const int4*>(arena + m_off);
        if constexpr (W > 0) k_many_w<K, W><<<total, BLOCK, 0, s>>>(a, m);
        else k_many<K><<<total, BLOCK, 0, s>>>(a, m);
    }});
}

// Per-chamber argument lists and block counts of one launch.
template <class K>
struct Batch {
    std::vector<typename K::Args> args;
    std::vector<int> blocks;
    void add(const typename K::Args& a, int nb) {
        args.push_back(a);
        blocks.push_back(nb);
    }
};

// Everything the device side of one world needs, by value.
// Sizes: `cap` is the reference capacity (logical slots, sorted positions);
// `pcap` = 2 x cap is the physical extent, room for a tick in which every
// cell divides on top of the tombstones left since the last re-sort.
template <class P>
struct WorldDev {
    Cells<P> c;          // the live SoA, physical order (pcap entries)
    Cells<P> alt;        // re-sort target (pcap entries)
    Counters* ctr;
    int32_t capacity;
    int32_t pcap;
    int32_t resort_period;  // re-sort at least every this many ticks

    Grid g[ASTRO_FIELD_COUNT];
    bool on[ASTRO_FIELD_COUNT];

    HashGeo hg;
    uint32_t* keys;         // [pcap] bucket of each physical entry
    uint32_t* rin;          // [pcap] arrival rank inside the bucket; reused as
                            //        the CO2 voxel of each cell during lifecycle
    float*    co2_avail;    // [pcap] CO2 concentration each cell's demand saw
    uint32_t* bcount;       // [nb + 1] bucket populations, zero between builds
    int32_t*  bstart;       // [nb + 2] dense exclusive prefix of bcount
    Rec<P>*   trec;         // [cap] records in arrival order
    uint32_t* tphys;        // [cap] physical index of each arrival-order record
    Rec<P>*   rec;          // [cap] records in stable bucket order
    uint32_t* vals_sorted;  // [cap] reference slot of each sorted record
    uint32_t* sphys;        // [cap] physical index of each sorted record
    int32_t*  spos;         // [pcap] sorted position of each physical entry
    Force<P>* fs;           // [cap] force in sorted order, at the layout's width
    double*   dirs;         // [cap] directional irradiance in sorted order

    int32_t*  bsum;         // [tiles(nb + 1)] per-scan-tile totals of bcount
    int32_t*  dflag;        // [cap] divides flag by reference slot
    int32_t*  prefix;       // [cap] exclusive prefix of dflag
    int32_t*  dsum;         // [tiles(cap)] per-scan-tile totals of dflag
    int32_t*  drop;         // [cap] compaction drop flag by reference slot
    int32_t*  dprefix;      // [cap] exclusive prefix of drop
    int32_t*  xsum;         // [tiles(cap)] per-scan-tile totals of drop
    uint32_t* dlist;        // [cap] dropped reference slots, ascending
    uint32_t* co2_demand;   // [CO2 grid] cells booking CO2 demand per voxel, zero between ticks
    uint32_t* tcount;       // [T grid] live cells per voxel during the thermal step, else 0
    uint8_t*  talone;       // [pcap] each entry's voxel this tick: 1 alone, 2 shared, 0 not live
    unsigned long long* stats;        // StatsAccum words (see launch_stats)
    unsigned long long* flash;        // [4] spin-drive audit: impulse x, y, z, energy
};

// Scalars of one tick, computed on the host from the world description.
struct TickParams {
    double dt, dt_bio;
    double chamber_w, chamber_h, chamber_d;
    int    boundary_x, boundary_y, gravity_axis;
    int    thermal_noise, contact, adhesion, occlusion_exact;
    int    store_disposition, compaction;
    double k_contact;
    double petrova_max_power, petrova_flash_power, quota, division_offset;
    // light
    float  source, ambient_irr;
    double ldx, ldy, ldz;
    int    sweep_axis, sweep_sign;
    double face_area;
    // thermal
    double T_cap;
    // lifecycle
    double co2_gv;
    // per-cell streams: Philox key (the world seed) when philox != 0
    uint64_t rng_seed;
    int    philox;
};

// The device views and tick scalars of the chambers one launch covers.
template <class P> using Views = std::span<const WorldDev<P>>;
using Ticks = std::span<const TickParams>;

// A span of one: the argument of a launcher for a single world.
template <class T> std::span<const T> one(const T& x) { return std::span<const T>(&x, 1); }

// FTCS coefficients of one field for one substep.
struct Ftcs {
    bool  square2d;
    float coeff;                 // square 2D
    float cx, cy, cz;            // general
    int   bc;
    float amb, rx, ry, rz;
};

// Launchers of the tick take the views (and tick scalars) of every chamber
// the launch covers: one for a world, all of an ensemble's chambers under a
// recorder. Per-chamber host decisions (which kernels run, substep counts)
// must be the same across the chambers of one launch.

// ---- hash (kernels_hash.hip) ----
// thermal_count: the key pass also counts the occupied entries per
// temperature voxel into tcount, as launch_thermal_count would.
template <class P> void launch_hash_build(const Launch& L, Views<P> ws, bool count_tick, bool thermal_count = false);
// The contact half of forces_kernel over the sorted records, adding to fs.
template <class P> void launch_contact_walk(const Launch& L, Views<P> ws, Ticks ts);
// A world with crowded[i] set has each cell's neighbourhood walked by a
// 16-lane group (one lane per row) instead of one lane: the same result,
// faster where cells screen many neighbours each.
template <class P> void launch_feed_walk(const Launch& L, Views<P> ws, Ticks ts, const std::vector<char>& crowded);
// Re-sorts the SoA into bucket order when due (device-side decision).
template <class P> void launch_resort(const Launch& L, Views<P> ws);

// ---- cells (kernels_cells.hip) ----
struct SpawnArgs {
    int32_t count, placement, charge_dist, awake;
    int32_t philox;              // draw from the Philox stream keyed by the spawn seed
    double place_x, place_y, place_radius, charge_a, charge_b;
    double chamber_w, chamber_h, chamber_d;
};
template <class P> void launch_spawn(const Launch& L, const WorldDev<P>& w, const SpawnArgs& a,
                                     int32_t base, int32_t lbase, uint64_t first_id, uint64_t seed);
// The thermostat's deposit paths (THERMAL_* in kernels_cells.hip): 0 atomic;
// 1 direct, first substep of the tick (after launch_thermal_count); 2 direct.
// `substep` counts the tick's substeps from 0; `last` marks the final one.
// `stamp`: the last substep also stamps the occluders of the chambers with a
// positive light source, as launch_stamp would after it.
template <class P> void launch_thermal_count(const Launch& L, Views<P> ws);
template <class P> void launch_thermal_substep(const Launch& L, Views<P> ws, Ticks ts, double dt_sub,
                                               bool update_rec, int mode, int substep, bool last,
                                               bool stamp = false);
template <class P> void launch_apply_awake(const Launch& L, Views<P> ws, int field, int mode, int substep);
template <class P> void launch_feed_taxis(const Launch& L, Views<P> ws, Ticks ts, bool feed, bool taxis);
// Feeding/taxis (optional), field sample + body force, then the contact walk:
// leaves each cell's total force in fs (sorted order).
template <class P> void launch_cell_forces(const Launch& L, Views<P> ws, Ticks ts, bool feed, bool taxis);
// flash_step (src/sim/flash.cu:59-67): the armed spin-drive flash in place
// of taxis.
template <class P> void launch_flash(const Launch& L, Views<P> ws, Ticks ts);
// integrate_kernel; with `demand` it also books each moved cell's CO2 demand
// (the lifecycle's first kernel; the CO2 field must have diffused already).
template <class P> void launch_integrate(const Launch& L, Views<P> ws, Ticks ts, bool demand);
// lifecycle_step; `demand_booked` when the integrate pass booked the demand.
template <class P> void launch_lifecycle(const Launch& L, Views<P> ws, Ticks ts, bool demand_booked);
// cell_store_compact outside the lifecycle: every dead or empty slot goes.
template <class P> void launch_compact(const Launch& L, const WorldDev<P>& w);
// Clears the per-slot lifecycle flags and their tile totals (the tick does
// this in its key pass).
template <class P> void launch_life_clear(const Launch& L, const WorldDev<P>& w);
template <class P> void launch_stats_cells(const Launch& L, const WorldDev<P>& w);

// ---- fields (kernels_fields.hip) ----
void launch_ftcs(const Launch& L, std::span<const Grid> gs, std::span<const Ftcs> fs, const char* name);
template <class P> void launch_stamp(const Launch& L, Views<P> ws);
void launch_sweep(const Launch& L, std::span<const Grid> irr, Ticks ts);
void launch_fill(const Launch& L, float* v, size_t n, float value);
void launch_brush(const Launch& L, const Grid& g, float cx, float cy, float radius, float amount);
void launch_stats_fields(const Launch& L, const Grid& T, const Grid& C, const Grid& N, bool c_on,
                         bool n_on, unsigned long long* stats);

// ---- Taumoeba (kernels_tau.hip) ----
// Device-resident bookkeeping of the predator store, so a tick never reads a
// count back (the reference syncs twice per predation step).
struct TauCtr {
    int32_t count, capacity;
    int32_t births, deads;      // this tick's (clamped) divisions, dead slots before them
    uint64_t next_id;
};

// The Taumoeba SoA (src/sim/predation.cuh:85-126), FP64 kinematics in either
// cell layout, plus the scan and gather scratch of its lifecycle.
struct TauDev {
    uint64_t* id;
    uint32_t* flags;
    double *x, *y, *z, *vx, *vy, *vz;
    float *dir_x, *dir_y, *dir_z;
    double *biomass, *prey_biomass;
    float *digest_timer, *tolerance;
    uint32_t* generation;
    uint64_t* rng_state;
    float *density_ema, *run_timer;
    int32_t* target;             // reference slot of the cell claimed this tick, -1 none
    TauCtr* ctr;
    int32_t* prey_rec;           // [capacity] sorted hash record of that cell
    int32_t capacity;
    int32_t cell_capacity;       // entries of `claim`
    int32_t* mark;               // [capacity] divides / survives flags
    int32_t* prefix;             // [capacity] their exclusive prefix
    int32_t* src;                // [capacity] compaction source slot of each output slot
    unsigned long long* scratch; // [capacity] gather buffer, 8 bytes a slot
    unsigned long long* claim;   // [cell capacity] per cell slot; all-ones between ticks
};

struct TauTick {
    double dt, dt_bio, ambient_temp;
    double chamber_w, chamber_h, chamber_d;
    int boundary_x, boundary_y, thermal_noise, store_disposition, compaction;
};

// predation_step (src/sim/predation.cu:435-490): hunt and claim, resolve,
// release the claims, nitrogen, then division and compaction in one block.
template <class P> void launch_predation(const Launch& L, const WorldDev<P>& w, const TauDev& t, const TauTick& tt);
// spawn_kernel (src/sim/predation.cu:33-76) for slots [base, base + count).
void launch_tau_spawn(const Launch& L, const TauDev& t, int32_t count, int32_t base, uint64_t first_id,
                      uint64_t seed, double chamber_w, double chamber_h, double chamber_d);
void launch_fill_u64(const Launch& L, unsigned long long* v, size_t n, unsigned long long value);
// stats_taumoeba_kernel (src/sim/stats.cu:65-74) into the stats words.
void launch_tau_stats(const Launch& L, const TauDev& t, unsigned long long* stats);

// StatsAccum word layout (src/sim/world.cuh:39-45), all fixed point.
enum StatsWord {
    SW_ENERGY = 0, SW_TEMP_CELL, SW_TEMP_MEDIUM, SW_CO2, SW_N2, SW_MAX_TEMP,
    SW_N_ALIVE, SW_N_DEAD, SW_N_AWAKE, SW_N_DORMANT, SW_N_TAU, SW_TAU_TOL, SW_COUNT
};

}  // namespace astro_dev
