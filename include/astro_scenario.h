/* astro_scenario.h -- scenarios over the MI355X world: the parsed form of
 * the scenario JSON files (scenarios/NAME.json), their instantiation, the
 * per-tick driving script and the acceptance runner.
 *
 * Stands in for the reference's scenario layer (src/sim/scenario.h,
 * src/sim/scenario.cpp), its acceptance evaluation (src/sim/accept.h,
 * src/sim/accept.cpp) and the headless --assert runner
 * (tools/headless.cpp:179-250). Host code only; it drives the world through
 * the C ABI of astro_hot.h and never touches device memory itself.
 *
 * Not carried: the microscope view (`scope`, render-only).
 */
#ifndef ASTRO_SCENARIO_H
#define ASTRO_SCENARIO_H

#include <stddef.h>
#include <stdint.h>

#include "astro_hot.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ASTRO_SCENARIO_MAX_LIGHTS      8
#define ASTRO_SCENARIO_MAX_POPULATIONS 8
#define ASTRO_SCENARIO_MAX_ACCEPT      16
#define ASTRO_SCENARIO_MAX_OVERRIDES   32
#define ASTRO_SCENARIO_MAX_STIMULI     16

/* OrganismKind, DensityModel, StimulusKind (contracts/scenario_v3.h:29-89),
 * Metric and CompareOp (contracts/telemetry_v1.h:43-58). */
enum { ASTRO_ORGANISM_ASTROPHAGE = 0, ASTRO_ORGANISM_TAUMOEBA = 1 };
enum { ASTRO_DENSITY_CANON_MASS = 0, ASTRO_DENSITY_WATER = 1 };
enum { ASTRO_STIM_HEAT = 0, ASTRO_STIM_CHILL = 1, ASTRO_STIM_INJECT_CO2 = 2, ASTRO_STIM_INJECT_N2 = 3,
       ASTRO_STIM_SET_LIGHT = 4, ASTRO_STIM_SET_N2 = 5, ASTRO_STIM_SEED_CELLS = 6, ASTRO_STIM_FLASH = 7 };
enum {
    ASTRO_METRIC_AWAKE_FRACTION = 0, ASTRO_METRIC_MEDIUM_TEMP_MEAN, ASTRO_METRIC_MEDIUM_TEMP_MAX,
    ASTRO_METRIC_BOIL_EVENT_COUNT, ASTRO_METRIC_POPULATION_LIVE, ASTRO_METRIC_POPULATION_DEAD,
    ASTRO_METRIC_MEAN_CHARGE, ASTRO_METRIC_TOTAL_ENERGY_J, ASTRO_METRIC_MEAN_TAU_TOLERANCE,
    ASTRO_METRIC_MAX_TAU_TOLERANCE, ASTRO_METRIC_CHARGE_DEPTH_CORRELATION,
    ASTRO_METRIC_CHARGE_HEIGHT_CORRELATION, ASTRO_METRIC_RISE_VELOCITY_EMPTY,
    ASTRO_METRIC_FALL_VELOCITY_FULL, ASTRO_METRIC_DOUBLING_TIME_S, ASTRO_METRIC_IMPULSE_PER_CYCLE,
    ASTRO_METRIC_COUNT
};
enum { ASTRO_OP_EQ = 0, ASTRO_OP_NE, ASTRO_OP_LT, ASTRO_OP_LE, ASTRO_OP_GT, ASTRO_OP_GE, ASTRO_OP_APPROX };

/* contract::LightSource (contracts/fields_v1.h:106-111). */
typedef struct astro_light_source {
    float dir_x, dir_y, irradiance, wavelength;
    int32_t enabled;
} astro_light_source;

/* contract::PopulationSpec (contracts/scenario_v3.h:45-53). */
typedef struct astro_population_spec {
    int32_t kind, count, placement, charge_dist;
    double place_x, place_y, place_radius;
    double charge_a, charge_b;
    int32_t awake, _pad;
} astro_population_spec;

/* contract::Stimulus (contracts/scenario_v3.h:91-97): active while
 * t0_s <= sim_time < t1_s (t1_s <= t0_s: open-ended); strength lerped v0 -> v1
 * across the window; radius <= 0 is chamber-global. */
typedef struct astro_stimulus {
    double t0_s, t1_s, x, y, radius, v0, v1;
    int32_t kind, _pad;
} astro_stimulus;

/* contract::AcceptCheck (contracts/telemetry_v1.h:60-67). */
typedef struct astro_accept_check {
    int32_t metric, op;
    double value, tol, after_s;
    int32_t tol_absolute, _pad;
} astro_accept_check;

typedef struct astro_param_override {
    char key[48];
    double value;
} astro_param_override;

/* contract::Scenario (contracts/scenario_v3.h:104-146) without `scope`. */
typedef struct astro_scenario {
    int32_t schema;
    char id[64];
    char title[128];
    char blurb[256];
    uint64_t seed;
    double chamber_w, chamber_h, chamber_d;
    int32_t boundary_x, boundary_y;
    double temp_init, co2_init, n2_init, ambient_temp;
    int32_t thermal_bc, density_model, store_disposition, gravity_axis;
    int32_t thermal_active, compaction, tau_compaction;
    float ambient_irradiance;
    astro_light_source lights[ASTRO_SCENARIO_MAX_LIGHTS];
    astro_population_spec populations[ASTRO_SCENARIO_MAX_POPULATIONS];
    int32_t population_count;
    int32_t clock;
    double physics_rate, biology_rate;
    uint32_t tools;
    int32_t override_count;
    astro_param_override param_overrides[ASTRO_SCENARIO_MAX_OVERRIDES];
    astro_stimulus drive[ASTRO_SCENARIO_MAX_STIMULI];
    int32_t drive_count;
    char objective_text[256];
    int32_t accept_count;
    astro_accept_check accept[ASTRO_SCENARIO_MAX_ACCEPT];
    double run_duration_s;
} astro_scenario;

/* scenario_parse (src/sim/scenario.cpp:139-318): jsonc text (with // and
 * block comments). Not an object, an unknown stimulus kind or an unknown
 * accept metric: ASTRO_INVALID_ARGUMENT. Unknown enum strings fall back to
 * the documented defaults, as the reference's do. */
int32_t astro_scenario_parse(const char* text, size_t len, astro_scenario* out);
/* scenario_load (src/sim/scenario.cpp:320-325). */
int32_t astro_scenario_load(const char* path, astro_scenario* out);
/* scenario_capacities (src/sim/scenario.cpp:327-341). */
void astro_scenario_capacities(const astro_scenario* s, int32_t* cell_capacity, int32_t* tau_capacity);
/* scenario_instantiate (src/sim/scenario.cpp:343-397): creates *out with the
 * scenario's chamber, capacity, seed, medium, motion switches, populations,
 * light and clock. `base` (NULL: the defaults) supplies what the scenario does
 * not say: grid resolutions, precision, hash cell. */
int32_t astro_scenario_instantiate(const astro_scenario* s, const astro_world_desc* base, astro_world** out);
/* scenario_apply_drive (src/sim/scenario.cpp:399-468) for the world's current
 * tick: call once before each astro_world_step(w, 1). */
int32_t astro_scenario_apply_drive(astro_world* w, const astro_scenario* s);

/* One evaluated accept check. */
typedef struct astro_accept_result {
    int32_t metric, op;
    double measured, value, tol, after_s;
    int32_t tol_absolute, pass;
} astro_accept_result;

/* Totals of one acceptance run. */
typedef struct astro_assert_summary {
    int64_t ticks;          /* ticks stepped */
    double sim_time_s;      /* at the end */
    int32_t checks, passed;
    int32_t all_pass;       /* 1: ACCEPT (also for a scenario with no checks) */
    int32_t samples;        /* stats samples taken */
} astro_assert_summary;

/* run_scenario_assert (tools/headless.cpp:179-250): runs the horizon (or
 * `ticks_override` > 0 ticks), applying the drive every tick and sampling
 * stats every max(1, ticks / 2000) ticks, then measures every check.
 * `csv_path` (may be NULL) receives the telemetry rows
 * (tools/headless.cpp:157-173). `results` receives up to `cap` checks. */
int32_t astro_scenario_run_assert(const astro_scenario* s, astro_world* w, int64_t ticks_override,
                                  const char* csv_path, astro_accept_result* results, int32_t cap,
                                  astro_assert_summary* summary);
/* accept_eval (src/sim/accept.cpp:208-221). */
int32_t astro_accept_eval(const astro_accept_check* c, double measured);
/* metric_name (src/sim/accept.cpp:223-244): the JSON name, "?" if unknown. */
const char* astro_metric_name(int32_t metric);

#ifdef __cplusplus
}
#endif

#endif /* ASTRO_SCENARIO_H */
