// kernels.h -- host-side launchers of the hot-path kernels.
//
// Each launcher enqueues its kernels on the world's stream and never
// synchronizes. Kernel grids are sized for the capacity; the live population
// count is read on the device, so a launch sequence can be captured once into
// a hipGraph and replayed.
#pragma once

#include <string>

#include "astro_dev.h"
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
    Profiler*   prof;   // null unless profiling
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
    Force*    fs;           // [cap] force in sorted order
    double*   dirs;         // [cap] directional irradiance in sorted order

    int32_t*  bsum;         // [tiles(nb + 1)] per-scan-tile totals of bcount
    int32_t*  dflag;        // [cap] divides flag by reference slot
    int32_t*  prefix;       // [cap] exclusive prefix of dflag
    int32_t*  dsum;         // [tiles(cap)] per-scan-tile totals of dflag
    int32_t*  drop;         // [cap] compaction drop flag by reference slot
    int32_t*  dprefix;      // [cap] exclusive prefix of drop
    int32_t*  xsum;         // [tiles(cap)] per-scan-tile totals of drop
    uint32_t* dlist;        // [cap] dropped reference slots, ascending
    unsigned long long* co2_demand;   // [CO2 grid] fixed-point demand
    unsigned long long* stats;        // StatsAccum words (see launch_stats)
};

// Scalars of one tick, computed on the host from the world description.
struct TickParams {
    double dt, dt_bio;
    double chamber_w, chamber_h, chamber_d;
    int    boundary_x, boundary_y, gravity_axis;
    int    thermal_noise, contact, adhesion, occlusion_exact;
    int    store_disposition, compaction;
    double k_contact;
    double petrova_max_power, quota, division_offset;
    // light
    float  source, ambient_irr;
    double ldx, ldy, ldz;
    int    sweep_axis, sweep_sign;
    double face_area;
    // thermal
    double T_cap;
    // lifecycle
    double co2_gv;
};

// FTCS coefficients of one field for one substep.
struct Ftcs {
    bool  square2d;
    float coeff;                 // square 2D
    float cx, cy, cz;            // general
    int   bc;
    float amb, rx, ry, rz;
};

// ---- hash (kernels_hash.hip) ----
template <class P> void launch_hash_build(const Launch& L, const WorldDev<P>& w, bool count_tick);
// The contact half of forces_kernel over the sorted records, adding to fs.
template <class P> void launch_contact_walk(const Launch& L, const WorldDev<P>& w, const TickParams& t);
template <class P> void launch_feed_walk(const Launch& L, const WorldDev<P>& w, const TickParams& t);
// Re-sorts the SoA into bucket order when due (device-side decision).
template <class P> void launch_resort(const Launch& L, const WorldDev<P>& w);

// ---- cells (kernels_cells.hip) ----
struct SpawnArgs {
    int32_t count, placement, charge_dist, awake;
    double place_x, place_y, place_radius, charge_a, charge_b;
    double chamber_w, chamber_h, chamber_d;
};
template <class P> void launch_spawn(const Launch& L, const WorldDev<P>& w, const SpawnArgs& a,
                                     int32_t base, int32_t lbase, uint64_t first_id, uint64_t seed);
template <class P> void launch_thermal_substep(const Launch& L, const WorldDev<P>& w, double dt_sub,
                                               double cap, bool update_rec);
template <class P> void launch_apply_awake(const Launch& L, const WorldDev<P>& w, int field);
template <class P> void launch_feed_taxis(const Launch& L, const WorldDev<P>& w, const TickParams& t,
                                          bool feed, bool taxis);
// integrate_kernel; with `demand` it also books each moved cell's CO2 demand
// (the lifecycle's first kernel; the CO2 field must have diffused already).
// Feeding/taxis (optional), field sample + body force, then the contact walk:
// leaves each cell's total force in fs (sorted order).
template <class P> void launch_cell_forces(const Launch& L, const WorldDev<P>& w, const TickParams& t, bool feed,
                                           bool taxis);
template <class P> void launch_integrate(const Launch& L, const WorldDev<P>& w, const TickParams& t, bool demand);
// lifecycle_step; `demand_booked` when the integrate pass booked the demand.
template <class P> void launch_lifecycle(const Launch& L, const WorldDev<P>& w, const TickParams& t,
                                         bool demand_booked);
// cell_store_compact outside the lifecycle: every dead or empty slot goes.
template <class P> void launch_compact(const Launch& L, const WorldDev<P>& w);
// Clears the per-slot lifecycle flags and their tile totals (the tick does
// this in its key pass).
template <class P> void launch_life_clear(const Launch& L, const WorldDev<P>& w);
template <class P> void launch_stats_cells(const Launch& L, const WorldDev<P>& w);

// ---- fields (kernels_fields.hip) ----
void launch_ftcs(const Launch& L, const Grid& g, const Ftcs& f, const char* name);
template <class P> void launch_stamp(const Launch& L, const WorldDev<P>& w);
void launch_sweep(const Launch& L, const Grid& irr, const TickParams& t);
void launch_fill(const Launch& L, float* v, size_t n, float value);
void launch_stats_fields(const Launch& L, const Grid& T, const Grid& C, const Grid& N, bool c_on,
                         bool n_on, unsigned long long* stats);

// StatsAccum word layout (src/sim/world.cuh:39-45), all fixed point.
enum StatsWord {
    SW_ENERGY = 0, SW_TEMP_CELL, SW_TEMP_MEDIUM, SW_CO2, SW_N2, SW_MAX_TEMP,
    SW_N_ALIVE, SW_N_DEAD, SW_N_AWAKE, SW_N_DORMANT, SW_COUNT
};

}  // namespace astro_dev
