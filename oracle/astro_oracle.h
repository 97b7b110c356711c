/* astro_oracle.h -- CPU oracle of the hot path. TEST INFRASTRUCTURE ONLY.
 *
 * A single-threaded FP64 C++ restatement of the reference simulator's
 * world_step (src/sim/step.cu:109-184) and of every stage it calls. It exists
 * to check the HIP library: only tests, the smoke check and the CPU-baseline
 * leg of the benchmark load it. Nothing in the product links or calls it.
 *
 * The entry points mirror include/astro_hot.h one for one, with an "oracle_"
 * prefix, and take the same structs, so a test drives both through one
 * binding.
 */
#ifndef ASTRO_ORACLE_H
#define ASTRO_ORACLE_H

#include "astro_hot.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_world oracle_world;

void        oracle_world_desc_default(astro_world_desc* d);
int32_t     oracle_world_create(const astro_world_desc* d, oracle_world** out);
void        oracle_world_destroy(oracle_world* w);
/* Test-only, per world, off by default. On: an FP32-layout world rounds
 * positions and velocities to float where the device layout stores them
 * (integrator, division). */
int32_t     oracle_world_set_float_store(oracle_world* w, int32_t on);
const char* oracle_status_str(int32_t status);
const char* oracle_last_error(const oracle_world* w);
int32_t     oracle_world_desc_get(const oracle_world* w, astro_world_desc* out);

int32_t oracle_cells_spawn(oracle_world* w, const astro_spawn_params* p, uint64_t seed);
int32_t oracle_cells_upload(oracle_world* w, const astro_cell_soa* src, int32_t n);
int32_t oracle_cells_download(oracle_world* w, astro_cell_soa* dst, int32_t n, int32_t* n_out);
int32_t oracle_world_get_counters(oracle_world* w, astro_counters* out);
int32_t oracle_world_set_counters(oracle_world* w, const astro_counters* in);

int32_t oracle_field_dims(const oracle_world* w, int32_t field, astro_grid_dims* out,
                          double* dx, double* dy, double* dz);
int32_t oracle_field_upload(oracle_world* w, int32_t field, const float* v, size_t count);
int32_t oracle_field_download(oracle_world* w, int32_t field, float* v, size_t count);
int32_t oracle_field_deposit_download(oracle_world* w, int32_t field, int64_t* v, size_t count);

int32_t oracle_world_step(oracle_world* w, int32_t n_ticks);
int32_t oracle_world_sync(oracle_world* w);
int32_t oracle_world_stats(oracle_world* w, astro_stats* out);
int32_t oracle_world_state_hash(oracle_world* w, uint64_t* out);
int32_t oracle_world_set_clock(oracle_world* w, double physics_rate, double biology_rate);
int32_t oracle_world_set_clock_preset(oracle_world* w, int32_t preset, double custom_physics,
                                      double custom_biology);
int32_t oracle_world_set_light(oracle_world* w, float dir_x, float dir_y, float irradiance, float ambient,
                               int32_t enabled);
int32_t oracle_world_apply_brush(oracle_world* w, int32_t kind, double x, double y, double radius,
                                 double strength);
int32_t oracle_field_fill(oracle_world* w, int32_t field, float value);
int32_t oracle_world_set_flash(oracle_world* w, int32_t armed);
int32_t oracle_world_flash_audit(oracle_world* w, double* impulse_ns, double* discharged_j);
int32_t oracle_world_time(const oracle_world* w, uint64_t* tick, double* sim_time_s, double* physics_rate,
                          double* biology_rate);
int32_t oracle_taumoeba_spawn(oracle_world* w, int32_t count, uint64_t seed);
int32_t oracle_taumoeba_upload(oracle_world* w, const astro_tau_soa* src, int32_t n);
int32_t oracle_taumoeba_download(oracle_world* w, astro_tau_soa* dst, int32_t n, int32_t* count_out);
int32_t oracle_taumoeba_counters(oracle_world* w, int32_t* count, int32_t* capacity, uint64_t* next_id);
int32_t oracle_stage_predation(oracle_world* w, double dt);

int32_t oracle_stage_hash_build(oracle_world* w);
int32_t oracle_hash_dims(const oracle_world* w, int32_t* nx, int32_t* ny, int32_t* nz,
                         double* cell_size);
int32_t oracle_hash_download(oracle_world* w, uint32_t* vals_sorted, int32_t n,
                             int32_t* bucket_start, int32_t* bucket_end, int32_t nb);
int32_t oracle_stage_thermal(oracle_world* w, double dt);
int32_t oracle_stage_emission(oracle_world* w, double dt);
int32_t oracle_stage_taxis(oracle_world* w, double dt);
int32_t oracle_stage_forces(oracle_world* w);
int32_t oracle_forces_download(oracle_world* w, double* fx, double* fy, double* fz, int32_t n);
int32_t oracle_stage_integrate(oracle_world* w, double dt);
int32_t oracle_stage_motion(oracle_world* w, double dt);
int32_t oracle_stage_slow_fields(oracle_world* w, double dt);
int32_t oracle_stage_diffuse(oracle_world* w, int32_t field, double dt);
int32_t oracle_stage_lifecycle(oracle_world* w, double dt);
int32_t oracle_stage_compact(oracle_world* w);

/* Threads of the parallel loops: the OpenMP team in liboracle_omp.so, else 1. */
int32_t oracle_threads(void);

/* Counter-based streams (include/astro_hot.h, ASTRO_RNG_PHILOX). */
int32_t oracle_philox_blocks(const uint32_t* words, int32_t n, uint32_t* out);
int32_t oracle_philox_stream(uint64_t seed, uint64_t id, uint64_t tick, int32_t stage, int32_t n,
                             uint32_t* out);

/* Building blocks exposed for known-answer tests. */
uint64_t oracle_splitmix64(uint64_t x);
void     oracle_pcg32_seed(uint64_t initstate, uint64_t initseq, uint64_t* state, uint64_t* inc);
uint32_t oracle_pcg32_next(uint64_t* state, uint64_t inc);
uint64_t oracle_cell_stream_inc(uint64_t id);
uint64_t oracle_cell_rng_init(uint64_t seed, uint64_t id);
uint64_t oracle_pcg_split(uint64_t parent_state, uint64_t daughter_id);
double   oracle_uniform01d(uint64_t* state, uint64_t inc);
void     oracle_gaussian_pair(uint64_t* state, uint64_t inc, double* z0, double* z1);
void     oracle_division_axis(uint64_t daughter_id, double* ax, double* ay, double* az);
double   oracle_tau_drag(double t_kelvin);
int32_t  oracle_tau_overlaps_cell(const double tau[3], const double cell[3]);
int32_t  oracle_tau_should_tumble(float signal, float* ema, float run_timer, double dt);
double   oracle_tau_n2_hazard(double n2_local, double tolerance);
double   oracle_tau_n2_death_prob(double hazard, double dt_bio);
float    oracle_tau_daughter_tolerance(float parent_tolerance, double mutation);
double   oracle_ou_position_shape(double x);
void     oracle_ou_coefficients(double mass, double gamma, double t_kelvin, double dt,
                                double out6[6]);
void     oracle_ou_advance_axis(double* pos, double* vel, double force, double gamma,
                                const double c6[6], double dt, int32_t thermal_noise,
                                double z1, double z2);
double   oracle_water_viscosity(double t);
double   oracle_disc_overlap_fraction(double p, double a);
double   oracle_shadow_fraction(const double self[3], const double other[3],
                                const double light[3]);
double   oracle_lumped_exchange_energy(double t_cell, double t_grid, double g, double c,
                                       double dt);
float    oracle_taxis_ema_update(float ema, double signal, double dt);
double   oracle_taxis_tumble_angle(uint64_t* state, uint64_t inc);
int64_t  oracle_to_fixed(double v, double scale);
uint64_t oracle_fnv1a64(const void* data, size_t n, uint64_t h);

#ifdef __cplusplus
}
#endif

#endif /* ASTRO_ORACLE_H */
