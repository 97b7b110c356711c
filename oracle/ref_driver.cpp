// ref_driver.cpp -- exports the reference simulator's own host-callable physics
// (its ASTRO_HD headers, compiled from /root/reference by oracle/Makefile) as
// plain C functions, so tests can pin the oracle against the reference's code
// bit for bit. TEST INFRASTRUCTURE: built into oracle/_ref/, never shipped,
// never linked by the product. Only headers whose kernels-free half compiles
// with g++ are used (ASTRO_HD expands to nothing outside nvcc,
// src/core/units.h:12-19).
#include <cstddef>
#include <cstdint>
#include <cstring>

#include "contracts/fields_v1.h"
#include "contracts/snapshot_v1.h"
#include "core/canon_generated.h"
#include "core/fixed_atomic.cuh"
#include "core/rng.cuh"
#include "fields/grid.cuh"
#include "sim/emission.cuh"
#include "sim/hash.cuh"
#include "sim/integrator.cuh"
#include "sim/lifecycle.cuh"
#include "sim/predation.cuh"
#include "sim/taxis.cuh"
#include "sim/thermal.cuh"

using namespace astro;
using namespace astro::sim;

extern "C" {

uint64_t ref_splitmix64(uint64_t x) { return splitmix64(x); }
void ref_pcg32_seed(uint64_t is, uint64_t iq, uint64_t* st, uint64_t* inc) {
    const Pcg32 r = pcg32_seed(is, iq);
    *st = r.state;
    *inc = r.inc;
}
uint32_t ref_pcg32_next(uint64_t* st, uint64_t inc) {
    Pcg32 r{*st, inc};
    const uint32_t v = pcg32_next(r);
    *st = r.state;
    return v;
}
uint64_t ref_cell_stream_inc(uint64_t id) { return cell_stream_inc(id); }
uint64_t ref_cell_rng_init(uint64_t seed, uint64_t id) { return cell_rng_init(seed, id); }
uint64_t ref_pcg_split(uint64_t p, uint64_t d) { return pcg_split(p, d); }
double ref_uniform01d(uint64_t* st, uint64_t inc) {
    Pcg32 r{*st, inc};
    const double v = uniform01d(r);
    *st = r.state;
    return v;
}
void ref_gaussian_pair(uint64_t* st, uint64_t inc, double* z0, double* z1) {
    Pcg32 r{*st, inc};
    gaussian_pair(r, *z0, *z1);
    *st = r.state;
}
void ref_division_axis(uint64_t did, double* ax, double* ay, double* az) { division_axis(did, *ax, *ay, *az); }
double ref_daughter_energy(double e) { return daughter_energy(e); }
double ref_co2_uptake_rate(double c) { return co2_uptake_rate(c); }
double ref_ou_position_shape(double x) { return ou_position_shape(x); }
void ref_ou_coefficients(double m, double g, double t, double dt, double o[6]) {
    const OuCoefficients c = ou_coefficients(m, g, t, dt);
    o[0] = c.e1; o[1] = c.one_minus_e1; o[2] = c.inv_beta;
    o[3] = c.sigma_rr; o[4] = c.sigma_vv; o[5] = c.sigma_rv;
}
void ref_ou_advance_axis(double* pos, double* vel, double force, double gamma, const double c6[6],
                         double dt, int32_t noise, double z1, double z2) {
    OuCoefficients c;
    c.e1 = c6[0]; c.one_minus_e1 = c6[1]; c.inv_beta = c6[2];
    c.sigma_rr = c6[3]; c.sigma_vv = c6[4]; c.sigma_rv = c6[5];
    ou_advance_axis(*pos, *vel, force, gamma, c, dt, noise != 0, z1, z2);
}
double ref_water_viscosity(double t) { return water_viscosity(t); }
double ref_drag_coefficient(double t) { return drag_coefficient(t); }
double ref_cell_mass(double b, double e) { return cell_mass(b, e); }
double ref_buoyant_weight(double m) { return buoyant_weight(m); }
void ref_contact_force(const double s[3], const double o[3], double k, double out[3]) {
    const Vec3 f = contact_force(Vec3{s[0], s[1], s[2]}, Vec3{o[0], o[1], o[2]}, k);
    out[0] = f.x; out[1] = f.y; out[2] = f.z;
}
int32_t ref_apply_boundary_axis(double* p, double* v, double half, double radius, int32_t mode) {
    return apply_boundary_axis(*p, *v, half, radius, static_cast<Boundary>(mode)) ? 1 : 0;
}
double ref_lumped_exchange_energy(double tc, double tg, double g, double c, double dt) {
    return lumped_exchange_energy(tc, tg, g, c, dt);
}
double ref_disc_overlap_fraction(double p, double a) { return disc_overlap_fraction(p, a); }
double ref_shadow_fraction(const double s[3], const double o[3], const double l[3]) {
    return shadow_fraction(Vec3{s[0], s[1], s[2]}, Vec3{o[0], o[1], o[2]}, Vec3{l[0], l[1], l[2]});
}
double ref_column_transmittance(double a, double f) { return column_transmittance(a, f); }
float ref_taxis_ema_update(float ema, double s, double dt) { return taxis_ema_update(ema, s, dt); }
int32_t ref_taxis_select_state(double c, double irr, double co2) {
    return static_cast<int32_t>(taxis_select_state(c, irr, co2));
}
double ref_taxis_tumble_angle(uint64_t* st, uint64_t inc) {
    Pcg32 r{*st, inc};
    const double v = taxis_tumble_angle(r);
    *st = r.state;
    return v;
}
void ref_taxis_tumble(const double h[3], uint64_t* st, uint64_t inc, double out[3]) {
    Pcg32 r{*st, inc};
    const Vec3 v = taxis_tumble(Vec3{h[0], h[1], h[2]}, r);
    *st = r.state;
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
}
double ref_taxis_emit_power(int32_t s, double e, double dt, double mx) {
    return taxis_emit_power(static_cast<TaxisState>(s), e, dt, mx);
}
int64_t ref_to_fixed(double v, double s) { return to_fixed(v, s); }
double ref_from_fixed(uint64_t raw, double s) { return from_fixed(raw, s); }
uint64_t ref_fnv1a64(const void* d, size_t n, uint64_t h) { return contract::fnv1a64(d, n, h); }
double ref_tau_drag(double t) { return tau_drag(t); }
int32_t ref_tau_overlaps_cell(const double a[3], const double b[3]) {
    return tau_overlaps_cell(Vec3{a[0], a[1], a[2]}, Vec3{b[0], b[1], b[2]}) ? 1 : 0;
}
int32_t ref_tau_should_tumble(float signal, float* ema, float run, double dt) {
    return tau_should_tumble(signal, *ema, run, dt) ? 1 : 0;
}
double ref_tau_n2_hazard(double n2, double tol) { return tau_n2_hazard(n2, tol); }
double ref_tau_n2_death_prob(double hazard, double dt_bio) { return tau_n2_death_prob(hazard, dt_bio); }
float ref_tau_daughter_tolerance(float parent, double mutation) { return tau_daughter_tolerance(parent, mutation); }

// The ASPH layout (contracts/snapshot_v1.h:25-55): header and override sizes,
// the offset of every header field in declaration order, the magic, the
// version and the hash's starting value (its hash of no bytes). Returns how
// many values there are; writes at most `cap`.
int32_t ref_snapshot_layout(int64_t* out, int32_t cap) {
    using H = contract::SnapshotHeader;
    const int64_t v[] = {
        sizeof(H), sizeof(contract::ParamOverride),
        offsetof(H, magic), offsetof(H, version), offsetof(H, seed), offsetof(H, tick), offsetof(H, sim_time_s),
        offsetof(H, cell_count), offsetof(H, cell_capacity), offsetof(H, taumoeba_count),
        offsetof(H, taumoeba_capacity), offsetof(H, n_temp), offsetof(H, n_co2), offsetof(H, n_n2),
        offsetof(H, chamber_w), offsetof(H, chamber_h), offsetof(H, chamber_d), offsetof(H, physics_rate),
        offsetof(H, biology_rate), offsetof(H, next_cell_id), offsetof(H, override_count),
        offsetof(H, non_canon_run), offsetof(H, scenario_id), offsetof(H, build_describe),
        offsetof(H, state_hash), contract::SNAPSHOT_MAGIC, contract::SNAPSHOT_CONTRACT_VERSION,
        static_cast<int64_t>(contract::fnv1a64(nullptr, 0)),
    };
    const int32_t n = static_cast<int32_t>(sizeof(v) / sizeof(v[0]));
    for (int32_t i = 0; i < n && i < cap; ++i) out[i] = v[i];
    return n;
}
int32_t ref_substeps_for_dt(double dx, double D, double dt) { return fields::substeps_for_dt(dx, D, dt); }

// grid_sample / grid_index_nearest on a host array (src/fields/grid.cuh:109-141).
float ref_grid_sample(float* value, int32_t n, float dx, float x, float y) {
    contract::FieldView f{};
    f.value = value; f.n = n; f.dx = dx;
    return fields::grid_sample(f, x, y);
}
int32_t ref_grid_index_nearest(int32_t n, float dx, float x, float y) {
    contract::FieldView f{};
    f.n = n; f.dx = dx;
    return fields::grid_index_nearest(f, x, y);
}

// hash_key through the reference's own HashView (src/sim/hash.cuh:63-78),
// with the view filled exactly as hash_view does (src/sim/hash.cu:90-104).
uint32_t ref_hash_key(int32_t nx, int32_t ny, int32_t nz, double cell, double x, double y, double z) {
    HashView h{};
    h.nx = nx; h.ny = ny; h.nz = nz;
    h.bucket_count = nx * ny * nz;
    h.cell_size = static_cast<float>(cell);
    h.inv_cell_size = static_cast<float>(1.0 / cell);
    h.origin_x = static_cast<float>(-0.5 * nx * cell);
    h.origin_y = static_cast<float>(-0.5 * ny * cell);
    h.origin_z = static_cast<float>(-0.5 * nz * cell);
    return hash_key(h, x, y, z);
}

// Named constants, for the constant-table check.
int32_t ref_constant(const char* name, double* out) {
    struct Kv { const char* k; double v; };
    static const Kv table[] = {
        {"C_LIGHT", canon::C_LIGHT}, {"K_BOLTZ", canon::K_BOLTZ}, {"G_ACCEL", canon::G_ACCEL},
        {"CELL_DIAMETER", canon::CELL_DIAMETER}, {"CELL_RADIUS", canon::CELL_RADIUS},
        {"CELL_VOLUME", canon::CELL_VOLUME}, {"CELL_CROSS_SECTION", canon::CELL_CROSS_SECTION},
        {"CELL_SURFACE_AREA", canon::CELL_SURFACE_AREA}, {"CELL_MASS_DRY", canon::CELL_MASS_DRY},
        {"CELL_TEMP_SETPOINT", canon::CELL_TEMP_SETPOINT}, {"CELL_ENERGY_MAX", canon::CELL_ENERGY_MAX},
        {"CELL_ALBEDO_DEAD", canon::CELL_ALBEDO_DEAD}, {"CELL_LETHAL_TEMP", canon::CELL_LETHAL_TEMP},
        {"PETROVA_MAX_POWER", canon::PETROVA_MAX_POWER}, {"PETROVA_FLASH_POWER", canon::PETROVA_FLASH_POWER},
        {"CO2_MASS_PER_DIVISION", canon::CO2_MASS_PER_DIVISION},
        {"CO2_DIFFUSIVITY_WATER", canon::CO2_DIFFUSIVITY_WATER},
        {"LIFE_DIVISION_ENERGY_COST", canon::LIFE_DIVISION_ENERGY_COST},
        {"LIFE_CO2_HALF_SATURATION", canon::LIFE_CO2_HALF_SATURATION},
        {"LIFE_CO2_UPTAKE_MAX", canon::LIFE_CO2_UPTAKE_MAX},
        {"WATER_DENSITY", canon::WATER_DENSITY}, {"WATER_CONDUCTIVITY", canon::WATER_CONDUCTIVITY},
        {"WATER_SPECIFIC_HEAT", canon::WATER_SPECIFIC_HEAT},
        {"WATER_THERMAL_DIFFUSIVITY", canon::WATER_THERMAL_DIFFUSIVITY},
        {"VF_A", canon::VF_A}, {"VF_B", canon::VF_B}, {"VF_C", canon::VF_C},
        {"AIR_CONVECTION_H", canon::AIR_CONVECTION_H},
        {"AMBIENT_TEMP_DEFAULT", canon::AMBIENT_TEMP_DEFAULT},
        {"CONDUCTION_COEFF", canon::CONDUCTION_COEFF}, {"CONTACT_STIFFNESS", canon::CONTACT_STIFFNESS},
        {"WALL_STICKINESS", canon::WALL_STICKINESS}, {"WALL_STUCK_DRAG_MULT", canon::WALL_STUCK_DRAG_MULT},
        {"WALL_RELEASE_FORCE", canon::WALL_RELEASE_FORCE}, {"HASH_CELL_FACTOR", canon::HASH_CELL_FACTOR},
        {"DT_PHYSICS", canon::DT_PHYSICS}, {"CLOCK_PHYSICS_MIN", canon::CLOCK_PHYSICS_MIN},
        {"CLOCK_PHYSICS_MAX", canon::CLOCK_PHYSICS_MAX}, {"CLOCK_BIOLOGY_MIN", canon::CLOCK_BIOLOGY_MIN},
        {"CLOCK_BIOLOGY_MAX", canon::CLOCK_BIOLOGY_MAX}, {"TAXIS_MEMORY_TIME", canon::TAXIS_MEMORY_TIME},
        {"CLOCK_REALTIME_PHYSICS", canon::CLOCK_REALTIME_PHYSICS},
        {"CLOCK_REALTIME_BIOLOGY", canon::CLOCK_REALTIME_BIOLOGY},
        {"CLOCK_MOTION_PHYSICS", canon::CLOCK_MOTION_PHYSICS}, {"CLOCK_MOTION_BIOLOGY", canon::CLOCK_MOTION_BIOLOGY},
        {"CLOCK_METABOLIC_PHYSICS", canon::CLOCK_METABOLIC_PHYSICS},
        {"CLOCK_METABOLIC_BIOLOGY", canon::CLOCK_METABOLIC_BIOLOGY},
        {"CLOCK_GENERATIONAL_PHYSICS", canon::CLOCK_GENERATIONAL_PHYSICS},
        {"CLOCK_GENERATIONAL_BIOLOGY", canon::CLOCK_GENERATIONAL_BIOLOGY},
        {"TAXIS_DARK_THRESHOLD", canon::TAXIS_DARK_THRESHOLD},
        {"TAXIS_SEEK_FEED_BELOW", canon::TAXIS_SEEK_FEED_BELOW},
        {"TAXIS_SEEK_BREED_ABOVE", canon::TAXIS_SEEK_BREED_ABOVE},
        {"TAXIS_TUMBLE_ANGLE_MEAN", canon::TAXIS_TUMBLE_ANGLE_MEAN},
        {"TAXIS_RUN_MAX", canon::TAXIS_RUN_MAX}, {"CHAMBER_W", canon::CHAMBER_W},
        {"CHAMBER_H", canon::CHAMBER_H}, {"CHAMBER_D", canon::CHAMBER_D},
        {"FIELD_N_TEMP", canon::FIELD_N_TEMP}, {"FIELD_N_CO2", canon::FIELD_N_CO2},
        {"FIELD_N_N2", canon::FIELD_N_N2}, {"FIELD_N_IRRAD", canon::FIELD_N_IRRAD},
        {"MAX_CELLS", canon::MAX_CELLS}, {"DEFAULT_CELLS", canon::DEFAULT_CELLS},
        {"MAX_TAUMOEBA", canon::MAX_TAUMOEBA}, {"DEFAULT_TAUMOEBA", canon::DEFAULT_TAUMOEBA},
        {"TAU_DIAMETER", canon::TAU_DIAMETER}, {"TAU_RADIUS", canon::TAU_RADIUS},
        {"TAU_VOLUME", canon::TAU_VOLUME}, {"TAU_MASS", canon::TAU_MASS}, {"TAU_MASS_DRY", canon::TAU_MASS_DRY},
        {"TAU_CRAWL_SPEED", canon::TAU_CRAWL_SPEED}, {"TAU_DRAG_20C", canon::TAU_DRAG_20C},
        {"TAU_CRAWL_THRUST", canon::TAU_CRAWL_THRUST}, {"TAU_DIGEST_TIME", canon::TAU_DIGEST_TIME},
        {"TAU_BIOMASS_YIELD", canon::TAU_BIOMASS_YIELD}, {"TAU_DIVIDE_BIOMASS", canon::TAU_DIVIDE_BIOMASS},
        {"TAU_N2_LETHAL_CONC", canon::TAU_N2_LETHAL_CONC}, {"TAU_N2_LETHAL_TIME", canon::TAU_N2_LETHAL_TIME},
        {"TAU_N2_HAZARD_RATE", canon::TAU_N2_HAZARD_RATE},
        {"TAU_N2_TOLERANCE_INIT", canon::TAU_N2_TOLERANCE_INIT},
        {"TAU_N2_TOLERANCE_K", canon::TAU_N2_TOLERANCE_K}, {"TAU_MUTATION_SIGMA", canon::TAU_MUTATION_SIGMA},
        {"DEPOSIT_SCALE_TEMPERATURE", contract::DEPOSIT_SCALE_TEMPERATURE},
        {"DEPOSIT_SCALE_CO2", contract::DEPOSIT_SCALE_CO2},
        {"DEPOSIT_SCALE_N2", contract::DEPOSIT_SCALE_N2},
        {"DEPOSIT_SCALE_STATS", contract::DEPOSIT_SCALE_STATS},
        {"DEPOSIT_SCALE_IRRADIANCE", contract::DEPOSIT_SCALE_IRRADIANCE},
        {"IRRADIANCE_NOMINAL_DIFFUSIVITY", contract::IRRADIANCE_NOMINAL_DIFFUSIVITY},
        {"CELL_FLAG_OCCUPIED", contract::CELL_FLAG_OCCUPIED}, {"CELL_FLAG_ALIVE", contract::CELL_FLAG_ALIVE},
        {"CELL_FLAG_AWAKE", contract::CELL_FLAG_AWAKE}, {"CELL_FLAG_STUCK", contract::CELL_FLAG_STUCK},
    };
    for (const Kv& kv : table) {
        if (std::strcmp(kv.k, name) == 0) { *out = kv.v; return 1; }
    }
    return 0;
}

}  // extern "C"
