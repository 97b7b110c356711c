/* astro_hot.h -- C ABI of the MI355X per-step hot path.
 *
 * This is the drop-in boundary for the reference simulator's world API
 * (src/sim/world.cuh:183-243), cell store (src/sim/cell_store.cuh:61-105),
 * spatial hash (src/sim/hash.cuh:50-53) and field grids
 * (src/fields/grid.cuh:87-98). Every entry point is plain C: opaque handles,
 * plain pointers and sizes, int32 status codes. Host buffers passed in are
 * caller-owned and are copied synchronously; the world owns all device memory,
 * allocated once at create (no allocation inside a step).
 *
 * Threading: calls on one world are serialized by the caller (one host thread
 * per world, as in the reference, docs/ARCHITECTURE.md:73-78). Distinct worlds
 * may be driven concurrently. Each world runs on its own HIP stream. Step calls
 * return once work is enqueued; download, stats and hash calls synchronize.
 *
 * Precision: ASTRO_FP64_CONTRACT keeps the reference's contract-v1 SoA
 * (f64 kinematics, contracts/cell_store_v1.h:42-78); ASTRO_FP32_HOT stores
 * position and velocity as f32 in device memory and computes the integrator's
 * drag, OU coefficients and Gaussian draws in f32 (displacement sums, energies
 * and budgets stay f64). Host-side views are always the f64 contract layout.
 */
#ifndef ASTRO_HOT_H
#define ASTRO_HOT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASTRO_ABI_VERSION 2

/* Status codes, in the order of astro::Status (src/core/result.h:12-23). */
enum {
    ASTRO_OK = 0,
    ASTRO_OUT_OF_MEMORY = 1,
    ASTRO_DEVICE_ERROR = 2,            /* reference: CudaError */
    ASTRO_INVALID_ARGUMENT = 3,
    ASTRO_NOT_FOUND = 4,
    ASTRO_PARSE_ERROR = 5,
    ASTRO_CONTRACT_VERSION_MISMATCH = 6,
    ASTRO_CAPACITY_EXCEEDED = 7,
    ASTRO_FILE_IO_ERROR = 8,
    ASTRO_UNSUPPORTED = 9
};

enum { ASTRO_FP64_CONTRACT = 0, ASTRO_FP32_HOT = 1 };

/* Per-cell random streams. ASTRO_RNG_PCG32 is the reference's generator
 * (src/core/rng.cuh:23-117): a 64-bit PCG32 state per cell in rng_state,
 * advanced by every draw. ASTRO_RNG_PHILOX is an addition: Philox4x32-10
 * (Salmon, Moraes, Dror and Shaw, SC'11), counter based, keyed by the seed,
 * with the counter (stage and draw, tick, cell id); a cell's draws in a tick
 * are a pure function of (seed, id, tick), so no state is read or written
 * per draw and rng_state is unused (0). Taumoeba keep PCG32 streams. */
enum { ASTRO_RNG_PCG32 = 0, ASTRO_RNG_PHILOX = 1 };

/* Boundary (src/sim/integrator.cuh:227), gravity axis (:175),
 * store disposition (contracts/scenario_v3.h:35). */
enum { ASTRO_BOUNDARY_REFLECTING = 0, ASTRO_BOUNDARY_PERIODIC = 1, ASTRO_BOUNDARY_ABSORBING = 2 };
enum { ASTRO_GRAVITY_Y = 0, ASTRO_GRAVITY_Z = 1 };
enum { ASTRO_STORE_VOID = 0, ASTRO_STORE_FLASH = 1, ASTRO_STORE_RETAIN = 2 };

/* Field ids and boundary conditions (contracts/fields_v1.h:98-110). */
enum { ASTRO_FIELD_TEMPERATURE = 0, ASTRO_FIELD_CO2 = 1, ASTRO_FIELD_N2 = 2,
       ASTRO_FIELD_IRRADIANCE = 3, ASTRO_FIELD_COUNT = 4 };
enum { ASTRO_BC_DIRICHLET = 0, ASTRO_BC_NEUMANN = 1, ASTRO_BC_ROBIN = 2 };

/* Spawn placement and charge distributions (src/sim/cell_store.cuh:15-16);
 * ASTRO_DIST_LOG_UNIFORM is an addition: charge = exp(U[ln a, ln b]). */
enum { ASTRO_PLACE_UNIFORM = 0, ASTRO_PLACE_GAUSSIAN = 1, ASTRO_PLACE_GRID = 2, ASTRO_PLACE_DISC = 3 };
enum { ASTRO_DIST_CONSTANT = 0, ASTRO_DIST_UNIFORM = 1, ASTRO_DIST_NORMAL = 2,
       ASTRO_DIST_LOG_UNIFORM = 3 };

/* One grid's resolution. nz == 1 is the reference's depth-averaged 2D grid
 * (src/fields/grid.cuh:72-85); nz > 1 resolves the chamber depth with a
 * 7-point stencil. All zero selects the reference's n x n x 1 for that field;
 * nx < 0 disables the field (N2 only). */
typedef struct astro_grid_dims {
    int32_t nx, ny, nz;
} astro_grid_dims;

/* Mirrors WorldDesc + MotionConfig + the World knobs the step reads
 * (src/sim/world.cuh:48-180). astro_world_desc_default() fills the
 * reference defaults. */
typedef struct astro_world_desc {
    int32_t  abi_version;        /* ASTRO_ABI_VERSION */
    int32_t  capacity;           /* 0 = ASTRO_DEFAULT_CELLS */
    uint64_t seed;
    double   chamber_w, chamber_h, chamber_d;
    int32_t  precision;          /* ASTRO_FP64_CONTRACT | ASTRO_FP32_HOT */
    int32_t  tau_capacity;       /* Taumoeba store capacity; 0 = ASTRO_DEFAULT_TAUMOEBA */
    double   hash_cell_size;     /* 0 = HASH_CELL_FACTOR * CELL_DIAMETER (22 um) */
    astro_grid_dims grid[ASTRO_FIELD_COUNT];  /* per field; zeros = reference n x n x 1 */
    int32_t  _pad1;
    /* MotionConfig */
    uint8_t  boundary_x, boundary_y, gravity_axis, thermal_noise;
    uint8_t  contact_enabled, adhesion_enabled, thermal_enabled, emission_enabled;
    uint8_t  taxis_enabled, occlusion_exact, store_disposition, compaction_enabled;
    uint8_t  lifecycle_enabled, tau_compaction_enabled;
    uint8_t  rng_kind;           /* ASTRO_RNG_PCG32 (default) | ASTRO_RNG_PHILOX */
    uint8_t  _pad2[5];
    double   ambient_temp;
    double   co2_init;           /* initial CO2 concentration [kg/m^3] */
    double   division_offset;    /* daughter offset along the axis; 0 = CELL_RADIUS */
    /* clock (src/sim/step.cu:133, :190-215) */
    double   physics_rate, biology_rate;
    /* curated overrides (src/sim/world.cuh:168-170); petrova_flash_power is
     * at the end of the struct */
    double   petrova_max_power;
    double   co2_mass_per_division;
    /* directional light (contracts/fields_v1.h:189-194) */
    float    light_dir_x, light_dir_y, light_irradiance, ambient_irradiance;
    uint8_t  light_enabled, _pad3[7];
    /* World::petrova_flash_power (src/sim/world.cuh:169): the spin-drive
     * flash's discharge rate [W], a snapshot override like the other two */
    double   petrova_flash_power;
} astro_world_desc;

/* SpawnParams (src/sim/cell_store.cuh:18-27). */
typedef struct astro_spawn_params {
    int32_t count;
    int32_t placement;
    double  place_x, place_y, place_radius;
    int32_t charge_dist;
    int32_t awake;
    double  charge_a, charge_b;
} astro_spawn_params;

/* Host-side view of the cell SoA in contract-v1 declaration order
 * (contracts/cell_store_v1.h:42-74). Each pointer addresses n elements. */
typedef struct astro_cell_soa {
    uint64_t* id;
    uint32_t* flags;
    uint8_t*  death_cause;
    double *x, *y, *z;
    double *vx, *vy, *vz;
    double*   energy;
    float*    temp_cell;
    float*    emit_power;
    float *dir_x, *dir_y, *dir_z;
    double*   biomass;
    double*   co2_held;
    float*    age_s;
    uint64_t* rng_state;
    float*    taxis_memory;
    float*    run_timer;
    float *t_local, *irradiance, *co2_local, *n2_local;
} astro_cell_soa;

/* Host bookkeeping the reference keeps in CellStore/World
 * (src/sim/cell_store.cuh:36-59, src/sim/world.cuh:111-129). */
typedef struct astro_counters {
    int32_t  count;              /* slots [0, count) may be occupied */
    int32_t  capacity;
    uint64_t next_id;
    uint64_t tick;
    double   sim_time_s;
    int64_t  divisions_total;
    int64_t  deaths_total;
    int32_t  dead_slots_prev;
    int32_t  _pad;
    int64_t  cell_steps;         /* sum over ticks of the count each tick began with */
} astro_counters;

/* contract::Stats (contracts/telemetry_v1.h:17-39). */
typedef struct astro_stats {
    uint64_t tick;
    double   sim_time_s;
    int32_t  n_live, n_dead, n_dormant, n_awake, n_taumoeba;
    int32_t  _pad0;
    double   mean_charge;
    double   total_energy_j;
    double   mean_temp_cell_k;
    double   mean_temp_medium_k;
    double   max_temp_medium_k;
    double   co2_total_kg;
    double   n2_total_kg;
    double   mean_tau_tolerance;
    int32_t  divisions_this_window;
    int32_t  deaths_this_window;
    int32_t  boil_events;
    int32_t  ignitions_this_window;
    double   physics_rate, biology_rate;
    uint8_t  non_canon_run, _pad1[7];
} astro_stats;

typedef struct astro_world astro_world;

/* ---- lifetime (src/sim/world.cuh:222-223) ---------------------------------- */
void        astro_world_desc_default(astro_world_desc* d);
int32_t     astro_world_create(const astro_world_desc* d, astro_world** out);
void        astro_world_destroy(astro_world* w);
const char* astro_status_str(int32_t status);
const char* astro_last_error(const astro_world* w);
int32_t     astro_world_desc_get(const astro_world* w, astro_world_desc* out);

/* ---- population (src/sim/cell_store.cuh:61-105) ---------------------------- */
/* cell_store_spawn: appends p->count cells seeded from (seed, id). */
int32_t astro_cells_spawn(astro_world* w, const astro_spawn_params* p, uint64_t seed);
/* Replaces slots [0, n) with the given cells and sets count = n. */
int32_t astro_cells_upload(astro_world* w, const astro_cell_soa* src, int32_t n);
/* Copies slots [0, min(count, n)) out; returns the number copied in *n_out. */
int32_t astro_cells_download(astro_world* w, astro_cell_soa* dst, int32_t n, int32_t* n_out);
int32_t astro_world_get_counters(astro_world* w, astro_counters* out);
int32_t astro_world_set_counters(astro_world* w, const astro_counters* in);

/* ---- fields (src/fields/grid.cuh:87-98) ------------------------------------ */
int32_t astro_field_dims(const astro_world* w, int32_t field, astro_grid_dims* out,
                         double* dx, double* dy, double* dz);
int32_t astro_field_upload(astro_world* w, int32_t field, const float* v, size_t count);
int32_t astro_field_download(astro_world* w, int32_t field, float* v, size_t count);
/* Fixed-point deposit accumulator of a field (contracts/fields_v1.h:166). */
int32_t astro_field_deposit_download(astro_world* w, int32_t field, int64_t* v, size_t count);

/* ---- the step (src/sim/step.cu:109-184) ------------------------------------ */
int32_t astro_world_step(astro_world* w, int32_t n_ticks);
int32_t astro_world_sync(astro_world* w);
/* world_stats (src/sim/stats.cu:102-153): runs the fixed-point reduction. */
int32_t astro_world_stats(astro_world* w, astro_stats* out);
/* snapshot_state_hash (src/sim/snapshot.cu:123-129): FNV-1a 64 over the
 * overrides, the contract-ordered SoA [0, count), the Taumoeba SoA, then the
 * T, CO2, N2 values. */
int32_t astro_world_state_hash(astro_world* w, uint64_t* out);
/* world_set_clock with explicit rates (src/sim/step.cu:190-215). */
int32_t astro_world_set_clock(astro_world* w, double physics_rate, double biology_rate);
/* How the thermostat folds a cell's exchange into the temperature field. No
 * counterpart in the reference, whose thermal_kernel always adds into the
 * fixed-point accumulator (src/sim/thermal.cu:19-71): ATOMIC does that for
 * every depositor; DIRECT counts the entries per voxel once per tick and lets
 * a cell alone in its voxel add its own exchange to the value; AUTO (the
 * default) takes DIRECT when most live cells are awake. The field is the same
 * bit for bit on every path. An ensemble follows its first chamber. */
enum { ASTRO_DEPOSIT_AUTO = 0, ASTRO_DEPOSIT_ATOMIC = 1, ASTRO_DEPOSIT_DIRECT = 2 };
int32_t astro_world_set_deposit_path(astro_world* w, int32_t path);

/* ---- stimuli a scenario applies between ticks ------------------------------ */
/* ClockPreset (contracts/scenario_v3.h:37) and BrushKind (src/sim/world.cuh). */
enum { ASTRO_CLOCK_REALTIME = 0, ASTRO_CLOCK_MOTION = 1, ASTRO_CLOCK_METABOLIC = 2,
       ASTRO_CLOCK_GENERATIONAL = 3, ASTRO_CLOCK_CUSTOM = 4 };
enum { ASTRO_BRUSH_HEAT = 0, ASTRO_BRUSH_CHILL = 1, ASTRO_BRUSH_INJECT_CO2 = 2,
       ASTRO_BRUSH_INJECT_N2 = 3 };
/* world_set_clock(World&, ClockPreset, custom_physics, custom_biology)
 * (src/sim/step.cu:190-215): a preset's rates, or the custom ones, clamped. */
int32_t astro_world_set_clock_preset(astro_world* w, int32_t preset, double custom_physics,
                                     double custom_biology);
/* World::light (the primary source) and World::ambient_irradiance
 * (src/sim/world.cuh, set by src/sim/scenario.cpp:394-395 and :431-434). */
int32_t astro_world_set_light(astro_world* w, float dir_x, float dir_y, float irradiance,
                              float ambient, int32_t enabled);
/* world_apply_brush (src/sim/step.cu:80-89): grid_brush's smooth disc
 * (src/fields/diffuse.cu:62-89) of `strength` on T (heat; chill subtracts),
 * CO2 or N2, centred at (x, y); radius <= 0 does nothing. A 3D grid takes the
 * same amount at every depth. Enqueued on the world's stream. */
int32_t astro_world_apply_brush(astro_world* w, int32_t kind, double x, double y, double radius,
                                double strength);
/* grid_fill (src/fields/diffuse.cu:128-134). */
int32_t astro_field_fill(astro_world* w, int32_t field, float value);
/* World::flash_active (src/sim/world.cuh:178): while armed, each tick runs
 * flash_step (src/sim/flash.cu:31-67) instead of taxis -- every awake
 * charged cell discharges at PETROVA_FLASH_POWER toward the objective. */
int32_t astro_world_set_flash(astro_world* w, int32_t armed);
/* world_flash_audit (src/sim/flash.cu:69-79): |impulse| [N s] and the
 * discharged energy [J] since the world was created. */
int32_t astro_world_flash_audit(astro_world* w, double* impulse_ns, double* discharged_j);
/* World::tick, sim_time_s and the clock rates, read on the host without
 * waiting for the device. */
int32_t astro_world_time(const astro_world* w, uint64_t* tick, double* sim_time_s,
                         double* physics_rate, double* biology_rate);

/* ---- Taumoeba, the predator (src/sim/predation.cuh) -------------------------- *
 * Its own SoA (FP64 kinematics in either cell layout) of desc.tau_capacity
 * slots. Once a world holds predators, each tick runs predation_step after
 * the slow fields and before the lifecycle (src/sim/step.cu:172-175): hunt
 * on the cell hash with lowest-id claims, engulf, N2 death, division with a
 * mutated tolerance, and, with desc.tau_compaction_enabled, compaction. */
/* TaumoebaStore's arrays in the reference's order (src/sim/predation.cuh:85-104,
 * serialized in that order by src/sim/snapshot.cu:59-67). Each pointer
 * addresses n elements. */
typedef struct astro_tau_soa {
    uint64_t* id;
    uint32_t* flags;
    double *x, *y, *z;
    double *vx, *vy, *vz;
    float *dir_x, *dir_y, *dir_z;
    double*   biomass;
    double*   prey_biomass;
    float*    digest_timer;
    float*    tolerance;
    uint32_t* generation;
    uint64_t* rng_state;
    float*    density_ema;
    float*    run_timer;
    int32_t*  target;
} astro_tau_soa;

/* taumoeba_spawn (src/sim/predation.cu:421-433): `count` predators placed
 * uniformly, streams from (splitmix64(seed), id). ASTRO_CAPACITY_EXCEEDED past
 * the store's capacity, with nothing spawned. */
int32_t astro_taumoeba_spawn(astro_world* w, int32_t count, uint64_t seed);
/* Slots [0, n) in; sets the count to n and next_id to max(id) + 1 (the
 * reference's snapshot rule, src/sim/snapshot.cu:239-252). */
int32_t astro_taumoeba_upload(astro_world* w, const astro_tau_soa* src, int32_t n);
/* Slots [0, min(n, count)) out; *count_out = the store's count. */
int32_t astro_taumoeba_download(astro_world* w, astro_tau_soa* dst, int32_t n, int32_t* count_out);
int32_t astro_taumoeba_counters(astro_world* w, int32_t* count, int32_t* capacity, uint64_t* next_id);
/* predation_step (src/sim/predation.cu:435-490) alone, on the hash of the
 * current population. */
int32_t astro_stage_predation(astro_world* w, double dt);

/* ---- ASPH snapshots (contracts/snapshot_v1.h, src/sim/snapshot.cu) ---------- *
 * The file is the contract's: a 256-byte header, then the parameter
 * overrides that differ from canon, the cell SoA [0, count) in contract
 * order, the Taumoeba SoA [0, taumoeba_count), and the T, CO2, N2 values;
 * the header's state_hash is FNV-1a 64 of that body (== astro_world_state_hash), started
 * from the contract's 1469598103934665603 (contracts/snapshot_v1.h:59). Header
 * field sizes follow the reference for its square 2D grids (n); a 3D grid
 * is written as minus its voxel count, an absent field as 0. */
/* snapshot_to_bytes (src/sim/snapshot.cu:131-160). `buf` NULL: only *n_out
 * (the size) is set; a `cap` below the size is ASTRO_CAPACITY_EXCEEDED. */
int32_t astro_snapshot_to_bytes(astro_world* w, const char* scenario_id, void* buf, size_t cap,
                                size_t* n_out);
/* snapshot_save (src/sim/snapshot.cu:162-170); ASTRO_FILE_IO_ERROR when the
 * file cannot be written. */
int32_t astro_snapshot_save(astro_world* w, const char* path, const char* scenario_id);
/* snapshot_from_bytes (src/sim/snapshot.cu:172-253): creates *out. The
 * motion configuration, grids and light are not in the file: they come from
 * `base` (NULL: the defaults), with the header's chamber, capacity, seed,
 * clock rates and overrides applied. Errors as the reference: shorter than a
 * header ASTRO_FILE_IO_ERROR; bad magic, hash mismatch (corrupt), field
 * resolution mismatch or truncated body ASTRO_INVALID_ARGUMENT; another
 * version ASTRO_CONTRACT_VERSION_MISMATCH. The Taumoeba next_id is
 * max(id) + 1, as the reference recovers it. */
int32_t astro_snapshot_from_bytes(const astro_world_desc* base, const void* data, size_t n,
                                  astro_world** out);
/* snapshot_load (src/sim/snapshot.cu:255-264). */
int32_t astro_snapshot_load(const astro_world_desc* base, const char* path, astro_world** out);

/* ---- the hash build without allocation ---------------------------------------
 * No counterpart in the reference. For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc. It is astro_stage_hash_build + astro_hash_download
 * (hash_build, src/sim/hash.cu:106-133) on caller-owned DEVICE arrays: the n
 * cells' positions (contract f64) and flags in, the stable order of slots by
 * bucket (free slots last, in slot order) and every bucket's [start, end)
 * out (start == end == 0 when empty), for a hash of nx x ny x nz buckets of
 * cell_size metres centred on the origin (as astro_hash_dims reports).
 * Nothing is allocated: every intermediate lives in one static device buffer
 * of ASTRO_HASH_NOALLOC_SCRATCH_BYTES (32 MiB) that each GPU the library runs
 * on holds until the process ends. n above ASTRO_HASH_NOALLOC_MAX_CELLS or
 * nx*ny*nz above ASTRO_HASH_NOALLOC_MAX_BUCKETS is ASTRO_CAPACITY_EXCEEDED.
 * Enqueued on `stream` (a hipStream_t; NULL = the null stream) without
 * waiting: the outputs are valid once the stream reaches this point. All
 * calls on one GPU share the buffer, so calls that overlap on a GPU (other
 * streams, other threads) race; serialize them. Should the build's scan
 * self-check fail, every bucket_start is -1. */
#define ASTRO_HASH_NOALLOC_MAX_CELLS   1048576
#define ASTRO_HASH_NOALLOC_MAX_BUCKETS 4194304
#define ASTRO_HASH_NOALLOC_SCRATCH_BYTES 33560320u
int32_t astro_hash_build_noalloc(const double* x, const double* y, const double* z, const uint32_t* flags,
                                 int32_t n, int32_t nx, int32_t ny, int32_t nz, double cell_size,
                                 uint32_t* vals_sorted, int32_t* bucket_start, int32_t* bucket_end,
                                 void* stream);

/* ---- stage entry points for kernel-level parity ---------------------------- */
/* hash_build (src/sim/hash.cu:106-133) on the current population. */
int32_t astro_stage_hash_build(astro_world* w);
int32_t astro_hash_dims(const astro_world* w, int32_t* nx, int32_t* ny, int32_t* nz,
                        double* cell_size);
/* vals_sorted[0,count), bucket_start/bucket_end[0,bucket_count) with the
 * reference's semantics (empty bucket: start == end == 0). */
int32_t astro_hash_download(astro_world* w, uint32_t* vals_sorted, int32_t n,
                            int32_t* bucket_start, int32_t* bucket_end, int32_t nb);
/* thermal_step (src/sim/thermal.cu:75-103). */
int32_t astro_stage_thermal(astro_world* w, double dt);
/* emission_step (src/sim/emission.cu:119-177), directional path. */
int32_t astro_stage_emission(astro_world* w, double dt);
/* taxis_step (src/sim/taxis.cu:82-87). */
int32_t astro_stage_taxis(astro_world* w, double dt);
/* field_sample + forces_kernel (src/sim/integrator.cu:20-76) into the force
 * buffer; astro_forces_download reads it back in slot order. */
int32_t astro_stage_forces(astro_world* w);
int32_t astro_forces_download(astro_world* w, double* fx, double* fy, double* fz, int32_t n);
/* integrate_kernel (src/sim/integrator.cu:79-150) from the force buffer. */
int32_t astro_stage_integrate(astro_world* w, double dt);
/* motion_step = sample + forces + integrate (src/sim/integrator.cu:154-177). */
int32_t astro_stage_motion(astro_world* w, double dt);
/* stages 7-8 for CO2 and N2 (src/sim/step.cu:168-171). */
int32_t astro_stage_slow_fields(astro_world* w, double dt);
/* one grid_diffuse (src/fields/diffuse.cu:159-186) of one field. */
int32_t astro_stage_diffuse(astro_world* w, int32_t field, double dt);
/* lifecycle_step (src/sim/lifecycle.cu:208-294) incl. the compaction trigger. */
int32_t astro_stage_lifecycle(astro_world* w, double dt);
/* cell_store_compact (src/sim/cell_store.cu:414-456). */
int32_t astro_stage_compact(astro_world* w);

/* ---- measurement ------------------------------------------------------------ */
/* Per-kernel HIP-event timing on the world's stream. While enabled, every
 * kernel launch of the step is bracketed by events and the step runs eagerly
 * (no graph). */
int32_t astro_profile_enable(astro_world* w, int32_t on);
/* Fills up to max entries: kernel name, launches, total milliseconds. */
int32_t astro_profile_read(astro_world* w, char* names, int32_t name_stride,
                           int64_t* launches, double* total_ms, int32_t max,
                           int32_t* n_out);
int32_t astro_profile_reset(astro_world* w);
/* Opaque hipStream_t of the world, for callers that order their own work. */
void*   astro_world_stream(astro_world* w);

/* ---- ensembles: independent chambers stepped together -------------------------
 * No counterpart in the reference, which steps one world per call
 * (src/sim/step.cu:109-184). An ensemble borrows n existing worlds (chambers
 * of any population and capacity) and advances all of them one tick at a
 * time: each stage is one batched kernel launch over every chamber, and a
 * step is one hipGraph launch. Every chamber's results are those of
 * astro_world_step on that world alone.
 *
 * The chambers must agree on what the host decides per tick: precision,
 * physics clock, enabled stages (thermal, light, taxis, lifecycle, contact),
 * which fields exist and each field's substep count. They may differ in cell
 * count, capacity, chamber size, light direction and irradiance, seeds and
 * field contents. Disagreement is reported by create and by step.
 *
 * The ensemble runs on a stream of its own, ordered after each chamber's
 * pending work at every step; each chamber's stream in turn waits for the
 * step, so the per-world calls (downloads, stats, state hash) see its result.
 * Worlds must outlive the ensemble and must not be stepped by other callers
 * while it uses them. */
typedef struct astro_ensemble astro_ensemble;
int32_t       astro_ensemble_create(astro_world* const* chambers, int32_t n, astro_ensemble** out);
void          astro_ensemble_destroy(astro_ensemble* e);
const char*   astro_ensemble_last_error(const astro_ensemble* e);
int32_t       astro_ensemble_size(const astro_ensemble* e);
astro_world*  astro_ensemble_chamber(astro_ensemble* e, int32_t i);
int32_t       astro_ensemble_step(astro_ensemble* e, int32_t n_ticks);
/* Waits for the ensemble's work and checks every chamber's device errors. */
int32_t       astro_ensemble_sync(astro_ensemble* e);
/* Picks launch shapes by measurement: for every stage, the total block count
 * a grid-stride stage is spread over across the chambers, and for the 3D FTCS
 * the lanes per x row, each the fastest of a few candidates. Every candidate
 * runs `ticks_per_trial` real, timed ticks (eleven trials in all), so the
 * chambers advance by 11 x ticks_per_trial ticks; launch shapes never change
 * results. The choice holds for later steps. */
int32_t       astro_ensemble_autotune(astro_ensemble* e, int32_t ticks_per_trial);
/* The current choice as a JSON object (NUL-terminated) in out[0, len). */
int32_t       astro_ensemble_tuning(const astro_ensemble* e, char* out, int32_t len);
/* Per-stage HIP-event timing of the batched launches; while enabled, steps
 * run eagerly (no graph). Same layout as astro_profile_read. */
int32_t       astro_ensemble_profile_enable(astro_ensemble* e, int32_t on);
int32_t       astro_ensemble_profile_read(astro_ensemble* e, char* names, int32_t name_stride,
                                          int64_t* launches, double* total_ms, int32_t max,
                                          int32_t* n_out);
int32_t       astro_ensemble_profile_reset(astro_ensemble* e);
void*         astro_ensemble_stream(astro_ensemble* e);

/* ---- counter-based streams (ASTRO_RNG_PHILOX) --------------------------------
 * Replace the per-cell PCG32 state (src/core/rng.cuh:23-117) in worlds
 * created with rng_kind = ASTRO_RNG_PHILOX. A cell's stream in one stage of
 * one tick is Philox4x32-10 keyed by the 64-bit seed (low word first) over the
 * counter (stage << 16 | block, tick, id low, id high), block = 0, 1, ...,
 * each block's four words drawn in order. Stages: 1 spawn (keyed by the
 * spawn seed, tick 0), 2 taxis tumble, 3 integrator (keyed by the world seed,
 * the tick in progress = ticks completed before it). The two calls below
 * evaluate the generator on the device, for known-answer checks. */
enum { ASTRO_RNG_STAGE_SPAWN = 1, ASTRO_RNG_STAGE_TAXIS = 2, ASTRO_RNG_STAGE_INTEGRATE = 3 };
/* out[4i .. 4i+3] = Philox4x32-10 of counter words[6i .. 6i+3] under key
 * words[6i+4], words[6i+5], for i < n. */
int32_t astro_philox_blocks(const uint32_t* words, int32_t n, uint32_t* out);
/* The first n 32-bit draws of the stream of cell `id` in `stage` of `tick`. */
int32_t astro_philox_stream(uint64_t seed, uint64_t id, uint64_t tick, int32_t stage, int32_t n,
                            uint32_t* out);

/* ---- learned field surrogate -------------------------------------------------
 * No counterpart in the reference. A 4-layer 3D convolutional operator (3x3x3
 * kernels, 32 channels, BF16 weights and activations, FP32 accumulation on
 * the matrix cores; hidden activations leaky ReLU with slope `leak`, linear
 * by default since the dynamics are linear in the fields) advances a temperature and a nutrient field by
 * `steps_per_call` FTCS substeps in one call. The dynamics it stands in for,
 * per substep: both fields take one 7-point FTCS substep with Neumann walls
 * (the reference's diffuse_kernel, src/fields/diffuse.cu:19-44, in 3D), then
 * T += dt * S and C -= dt * uptake * occupancy * C. It is trained on
 * trajectories of that plain path from random states (Gaussian warm/cool
 * spots, nutrient gradients, a clustered culture of `cells` cells giving
 * occupancy and heating). The plain path is also its judge and fallback. */
typedef struct astro_surrogate astro_surrogate;
typedef struct astro_surrogate_desc {
    int32_t  abi_version;
    int32_t  nx, ny, nz;            /* nx a multiple of 32 */
    double   dx, dy, dz;            /* m */
    double   d_temp, d_nutrient;    /* diffusivities, m^2/s */
    double   dt_sub;                /* s, stable for both fields */
    int32_t  steps_per_call;
    int32_t  cells;                 /* culture size of the training states */
    double   uptake;                /* nutrient uptake per occupant, 1/s */
    double   temp_ref, nutrient_ref;
    uint64_t seed;                  /* weight initialization */
    double   leak;                  /* hidden activation: leaky ReLU slope, 1 = linear (default) */
    /* gated calls: every judge_every-th call also runs the plain path from the
     * same start; a relative L2 error of the T increment above judge_tol, or a
     * heat-budget residual above budget_tol on any call, keeps the plain
     * result and runs the next `cooldown` calls plain. */
    int32_t  judge_every;
    double   judge_tol, budget_tol;
    int32_t  cooldown;
} astro_surrogate_desc;

enum { ASTRO_SURROGATE_PLAIN = 0, ASTRO_SURROGATE_NET = 1, ASTRO_SURROGATE_GATED = 2 };

typedef struct astro_surrogate_counters {
    int64_t calls, fallbacks, judged;
    double  last_judge_err;         /* relative L2 of the T increment, -1 if never judged */
    double  last_budget_residual;
    double  out_scale_temp, out_scale_nutrient;   /* K and kg/m^3 per unit output */
} astro_surrogate_counters;

void        astro_surrogate_desc_default(astro_surrogate_desc* d);   /* 128^3, 550x550x200 um, 16 substeps of 6.25 us */
int32_t     astro_surrogate_create(const astro_surrogate_desc* d, astro_surrogate** out);
void        astro_surrogate_destroy(astro_surrogate* s);
const char* astro_surrogate_last_error(const astro_surrogate* s);
/* Adam over `trajectories` states from `seed`, `epochs` passes, cosine-decayed
 * rate from lr; the mean loss of the last pass in *loss. */
int32_t     astro_surrogate_train(astro_surrogate* s, int32_t trajectories, int32_t epochs, float lr,
                                  uint64_t seed, double* loss);
/* A random state of the training distribution as the current state. */
int32_t     astro_surrogate_sample(astro_surrogate* s, uint64_t seed);
int32_t     astro_surrogate_set_state(astro_surrogate* s, const float* temp, const float* nutrient,
                                      const float* source, const float* occupancy, size_t n);
int32_t     astro_surrogate_get_state(astro_surrogate* s, float* temp, float* nutrient,
                                      float* source, float* occupancy, size_t n);
/* One call: mode PLAIN (steps_per_call FTCS substeps), NET, or GATED. */
int32_t     astro_surrogate_advance(astro_surrogate* s, int32_t mode, int32_t* used_net);
int32_t     astro_surrogate_counters_get(const astro_surrogate* s, astro_surrogate_counters* out);
/* Master weights of one layer ([27][32 out][cin] then [32] bias; cin = 8 for
 * layer 0, else 32): read, or write (write = 1, resets the Adam moments). */
int32_t     astro_surrogate_weights(astro_surrogate* s, int32_t layer, float* w, size_t n, int32_t write);
/* The network on normalized inputs x[n][8]: predictions pred[n][2]; with a
 * target[n][2], also the mean-squared-error gradient of every layer's
 * weights, layers in order (no update). */
int32_t     astro_surrogate_eval(astro_surrogate* s, const float* x, size_t n, const float* target,
                                 float* pred, float* grads, size_t ngrads);
void*       astro_surrogate_stream(astro_surrogate* s);

#ifdef __cplusplus
}
#endif

#endif /* ASTRO_HOT_H */
