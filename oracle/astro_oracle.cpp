// astro_oracle.cpp -- CPU oracle of the per-step hot path. TEST INFRASTRUCTURE.
//
// Single-threaded FP64 restatement of the reference simulator's world_step and
// its stages, written from the reference source text (file:line cited per
// function). Loaded only by the test suite, the smoke check and the benchmark's
// CPU-baseline leg; the HIP library never links or calls it.
//
// Geometry extension: a field with nz > 1 resolves the chamber depth. The
// reference grids are depth-averaged n x n (nz == 1), and for nz == 1 with a
// square n x n grid every formula below reduces to the reference's exactly.
//
// Compiled with -ffp-contract=off so the float and double expressions round
// exactly as written (the HIP library is compiled the same way).
#include "astro_oracle.h"

#include "astro_canon.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

// The all-cores CPU baseline builds this file with -fopenmp
// (liboracle_omp.so). Loops over cells or voxels that write only their own
// element, or add into int64 fixed-point accumulators (integer sums, so the
// order does not matter), then run on every core; all else stays sequential,
// and the results equal the single-threaded build's bit for bit. Without
// OpenMP the macros are plain loops and adds.
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#define PAR_FOR _Pragma("omp parallel for schedule(static)")
#define PAR_FOR2 _Pragma("omp parallel for collapse(2) schedule(static)")
#define FIXED_ADD(dst, v)                 \
    do {                                  \
        const uint64_t fixed_add_v = (v); \
        _Pragma("omp atomic")             \
        (dst) += fixed_add_v;             \
    } while (0)
#define STABLE_SORT __gnu_parallel::stable_sort
#else
#define PAR_FOR
#define PAR_FOR2
#define FIXED_ADD(dst, v) ((dst) += (v))
#define STABLE_SORT std::stable_sort
#endif

namespace {

// ---------------------------------------------------------------------------
// RNG: PCG32 with per-cell streams (src/core/rng.cuh:23-117)
// ---------------------------------------------------------------------------
inline uint64_t splitmix64(uint64_t x) {                     // rng.cuh:23-28
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// A cell's 32-bit draws: the reference's PCG32 state (state, inc), or, in a
// world created with rng_kind = ASTRO_RNG_PHILOX, a Philox4x32-10 counter
// stream (key, counter, the unread words of the last block).
struct Rng {
    uint64_t state = 0, inc = 0;
    bool philox = false;
    uint32_t key[2] = {0u, 0u};
    uint32_t ctr[4] = {0u, 0u, 0u, 0u};
    uint32_t buf[4] = {0u, 0u, 0u, 0u};
    int left = 0;
};

inline uint32_t pcg_step(Rng& r) {                           // rng.cuh:35-41
    const uint64_t old = r.state;
    r.state = old * 6364136223846793005ull + r.inc;
    const uint32_t xs = static_cast<uint32_t>(((old >> 18) ^ old) >> 27);
    const uint32_t rot = static_cast<uint32_t>(old >> 59);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, SC'11): ten rounds; each
// multiplies counter words 0 and 2 by M0 and M1 (32 x 32 -> 64 bits), and the
// new counter is (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0); the key advances by
// the Weyl constants W0, W1 before every round after the first.
inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint64_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = M0 * c[0], p1 = M1 * c[2];
        const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
        const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
}

// The stream of cell `id` in `stage` of `tick`: key = seed (low word first),
// counter = (stage << 16 | block, tick, id low, id high), blocks 0, 1, ...,
// each block's four words in order (include/astro_hot.h, ASTRO_RNG_PHILOX).
inline Rng philox_stream(uint64_t seed, uint32_t stage, uint64_t tick, uint64_t id) {
    Rng r;
    r.philox = true;
    r.key[0] = static_cast<uint32_t>(seed);
    r.key[1] = static_cast<uint32_t>(seed >> 32);
    r.ctr[0] = stage << 16;
    r.ctr[1] = static_cast<uint32_t>(tick);
    r.ctr[2] = static_cast<uint32_t>(id);
    r.ctr[3] = static_cast<uint32_t>(id >> 32);
    return r;
}

inline uint32_t philox_next(Rng& r) {
    if (r.left == 0) {
        for (int k = 0; k < 4; ++k) r.buf[k] = r.ctr[k];
        philox4x32_10(r.buf, r.key[0], r.key[1]);
        ++r.ctr[0];
        r.left = 4;
    }
    return r.buf[4 - r.left--];
}

inline uint32_t pcg_next(Rng& r) { return r.philox ? philox_next(r) : pcg_step(r); }

inline Rng pcg_seed(uint64_t initstate, uint64_t initseq) {  // rng.cuh:44-50
    Rng r{0ull, (initseq << 1u) | 1u};
    pcg_step(r);
    r.state += initstate;
    pcg_step(r);
    return r;
}

inline uint64_t stream_inc(uint64_t id) {                    // rng.cuh:54-56
    return (splitmix64(id ^ 0xA5A5A5A5DEADBEEFull) << 1u) | 1u;
}
inline Rng cell_rng(uint64_t st, uint64_t id) { return Rng{st, stream_inc(id)}; }
inline uint64_t cell_rng_init(uint64_t seed, uint64_t id) {  // rng.cuh:62-65
    return pcg_seed(seed ^ splitmix64(id), splitmix64(id + 1)).state;
}
inline uint64_t pcg_split(uint64_t parent, uint64_t did) {   // rng.cuh:69-71
    return splitmix64(parent ^ splitmix64(did * 0x2545F4914F6CDD1Dull));
}

inline double uniform01d(Rng& r) {                           // rng.cuh:80-84
    const uint64_t hi = static_cast<uint64_t>(pcg_next(r)) << 21;
    const uint64_t lo = static_cast<uint64_t>(pcg_next(r)) >> 11;
    return static_cast<double>((hi ^ lo) & ((1ull << 53) - 1)) * (1.0 / 9007199254740992.0);
}
inline double uniform_range(Rng& r, double lo, double hi) { return lo + (hi - lo) * uniform01d(r); }

inline void gaussian_pair(Rng& r, double& z0, double& z1) {  // rng.cuh:92-101
    double u1 = uniform01d(r);
    if (u1 < 1.0e-300) u1 = 1.0e-300;
    const double u2 = uniform01d(r);
    const double mag = std::sqrt(-2.0 * std::log(u1));
    const double ang = 6.283185307179586476925286766559 * u2;
    z0 = mag * std::cos(ang);
    z1 = mag * std::sin(ang);
}

// ---------------------------------------------------------------------------
// Fixed point (src/core/fixed_atomic.cuh:27-33)
// ---------------------------------------------------------------------------
inline int64_t to_fixed(double v, double scale) { return static_cast<int64_t>(std::llround(v * scale)); }
inline double from_fixed(uint64_t raw, double scale) {
    return static_cast<double>(static_cast<int64_t>(raw)) / scale;
}

// ---------------------------------------------------------------------------
// Small vector (src/core/vec.cuh)
// ---------------------------------------------------------------------------
struct V3 { double x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 normalize(V3 a) {
    const double l = std::sqrt(dot(a, a));
    return l > 0.0 ? a * (1.0 / l) : V3{0.0, 0.0, 1.0};
}
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ---------------------------------------------------------------------------
// Integrator physics (src/sim/integrator.cuh)
// ---------------------------------------------------------------------------
inline double water_viscosity(double t) {                    // integrator.cuh:29-31
    return ASTRO_VF_A * std::exp(ASTRO_VF_B / (t - ASTRO_VF_C));
}
inline double drag_coefficient(double t) { return 6.0 * ASTRO_PI * water_viscosity(t) * ASTRO_CELL_RADIUS; }
inline double cell_mass(double biomass, double energy) {     // integrator.cuh:42-44
    return biomass + energy / (ASTRO_C_LIGHT * ASTRO_C_LIGHT);
}
inline double buoyant_weight(double mass) {                  // integrator.cuh:51-53
    return -(mass - ASTRO_WATER_DENSITY * ASTRO_CELL_VOLUME) * ASTRO_G_ACCEL;
}

inline double ou_position_shape(double x) {                  // integrator.cuh:80-91
    if (x < 1.0e-2) {
        const double x3 = x * x * x;
        return x3 * (2.0 / 3.0 + x * (-0.5 + x * (7.0 / 30.0 + x * (-1.0 / 12.0))));
    }
    const double e = std::exp(-x);
    return 2.0 * x - 3.0 + 4.0 * e - e * e;
}

struct Ou { double e1, one_minus_e1, inv_beta, sigma_rr, sigma_vv, sigma_rv; };

inline Ou ou_coefficients(double mass, double gamma, double t, double dt) {  // :103-121
    const double beta = gamma / mass;
    const double x = beta * dt;
    Ou c;
    c.one_minus_e1 = -std::expm1(-x);
    c.e1 = 1.0 - c.one_minus_e1;
    c.inv_beta = 1.0 / beta;
    const double kT_over_m = ASTRO_K_BOLTZ * t / mass;
    const double var_vv = kT_over_m * (-std::expm1(-2.0 * x));
    const double var_rr = kT_over_m * c.inv_beta * c.inv_beta * ou_position_shape(x);
    c.sigma_vv = std::sqrt(var_vv > 0.0 ? var_vv : 0.0);
    c.sigma_rr = std::sqrt(var_rr > 0.0 ? var_rr : 0.0);
    c.sigma_rv = kT_over_m * c.inv_beta * c.one_minus_e1 * c.one_minus_e1;
    return c;
}

inline void ou_advance_axis(double& pos, double& vel, double force, double gamma,
                            const Ou& c, double dt, bool noise, double z1, double z2) {  // :128-154
    const double v_term = force / gamma;
    const double dpos = v_term * dt + (vel - v_term) * c.one_minus_e1 * c.inv_beta;
    const double dvel = (v_term - vel) * c.one_minus_e1;
    double npos = pos + dpos;
    double nvel = vel + dvel;
    if (noise && c.sigma_rr > 0.0) {
        const double a = c.sigma_rv / c.sigma_rr;
        const double b2 = c.sigma_vv * c.sigma_vv - a * a;
        const double b = std::sqrt(b2 > 0.0 ? b2 : 0.0);
        npos += c.sigma_rr * z1;
        nvel += a * z1 + b * z2;
    } else if (noise) {
        nvel += c.sigma_vv * z2;
    }
    pos = npos;
    vel = nvel;
}

inline V3 contact_force(V3 self, V3 other, double k) {       // integrator.cuh:210-219
    const V3 d = self - other;
    const double r2 = dot(d, d);
    const double two_a = ASTRO_CELL_DIAMETER;
    if (r2 >= two_a * two_a || r2 <= 0.0) return {0.0, 0.0, 0.0};
    const double r = std::sqrt(r2);
    const double overlap = two_a - r;
    return d * (k * overlap / r);
}

inline bool apply_boundary_axis(double& p, double& v, double half, double radius, int mode) {  // :233-254
    const double lo = -half + radius;
    const double hi = half - radius;
    if (p >= lo && p <= hi) return true;
    switch (mode) {
        case ASTRO_BOUNDARY_PERIODIC: {
            const double span = 2.0 * half;
            while (p < -half) p += span;
            while (p > half) p -= span;
            return true;
        }
        case ASTRO_BOUNDARY_ABSORBING: return false;
        default:
            p = (p < lo) ? lo : hi;
            v = 0.0;
            return true;
    }
}

// ---------------------------------------------------------------------------
// Thermal (src/sim/thermal.cuh)
// ---------------------------------------------------------------------------
inline double lumped_exchange_energy(double tc, double tg, double g, double cap, double dt) {  // :80-86
    if (cap <= 0.0 || g <= 0.0) return 0.0;
    const double relax = -std::expm1(-g * dt / cap);
    return cap * (tc - tg) * relax;
}

// ---------------------------------------------------------------------------
// Emission (src/sim/emission.cuh)
// ---------------------------------------------------------------------------
inline double disc_overlap_fraction(double p, double a) {    // emission.cuh:19-25
    if (p <= 0.0) return 1.0;
    if (p >= 2.0 * a) return 0.0;
    const double x = p / (2.0 * a);
    return (2.0 / ASTRO_PI) * (std::acos(x) - x * std::sqrt(1.0 - x * x));
}
inline double shadow_fraction(V3 self, V3 other, V3 light) {  // emission.cuh:30-37
    const V3 d = self - other;
    const double along = dot(d, light);
    if (along <= 0.0) return 0.0;
    const double perp2 = dot(d, d) - along * along;
    const double perp = perp2 > 0.0 ? std::sqrt(perp2) : 0.0;
    return disc_overlap_fraction(perp, ASTRO_CELL_RADIUS);
}
inline double column_transmittance(double blocked, double face) {  // emission.cuh:43-47
    if (face <= 0.0) return 1.0;
    const double t = 1.0 - blocked / face;
    return t > 0.0 ? t : 0.0;
}
inline double absorbed_power(double dir, double amb) {
    return dir * ASTRO_CELL_CROSS_SECTION + amb * ASTRO_CELL_SURFACE_AREA;
}

// ---------------------------------------------------------------------------
// Taxis (src/sim/taxis.cuh)
// ---------------------------------------------------------------------------
inline float taxis_ema_update(float ema, double signal, double dt) {  // taxis.cuh:31-34
    const double alpha = 1.0 - std::exp(-dt / ASTRO_TAXIS_MEMORY_TIME);
    return static_cast<float>(static_cast<double>(ema) + alpha * (signal - static_cast<double>(ema)));
}
enum TaxisState { IDLE = 0, FEED = 1, BREED = 2 };
inline int taxis_select_state(double charge, double irr, double co2) {  // taxis.cuh:55-62
    const bool dark = irr < ASTRO_TAXIS_DARK_THRESHOLD;
    const bool co2_here = co2 > 0.0;
    if (dark && !co2_here) return IDLE;
    if (!dark && charge < ASTRO_TAXIS_SEEK_FEED_BELOW) return FEED;
    if (charge > ASTRO_TAXIS_SEEK_BREED_ABOVE && co2_here) return BREED;
    return IDLE;
}
inline V3 any_perpendicular(V3 v) {                         // taxis.cuh:85-91
    const double ax = std::fabs(v.x), ay = std::fabs(v.y), az = std::fabs(v.z);
    const V3 axis = (ax <= ay && ax <= az) ? V3{1.0, 0.0, 0.0}
                  : (ay <= az)             ? V3{0.0, 1.0, 0.0}
                                           : V3{0.0, 0.0, 1.0};
    return normalize(cross(v, axis));
}
inline V3 rotate_by(V3 h, double theta, double az) {         // taxis.cuh:94-99
    const V3 u = any_perpendicular(h);
    const V3 w = cross(h, u);
    const V3 perp = u * std::cos(az) + w * std::sin(az);
    return normalize(h * std::cos(theta) + perp * std::sin(theta));
}
inline double taxis_tumble_angle(Rng& r) {                    // taxis.cuh:108-111
    const double theta = -ASTRO_TAXIS_TUMBLE_ANGLE_MEAN * std::log(1.0 - uniform01d(r));
    return theta > ASTRO_PI ? ASTRO_PI : theta;
}
inline V3 taxis_tumble(V3 heading, Rng& r) {                 // taxis.cuh:114-118
    const double theta = taxis_tumble_angle(r);
    const double azimuth = ASTRO_TWO_PI * uniform01d(r);
    return rotate_by(normalize(heading), theta, azimuth);
}
inline double taxis_emit_power(int s, double energy, double dt, double max_power) {  // :136-141
    if (s == IDLE || energy <= 0.0) return 0.0;
    const double avail = dt > 0.0 ? energy / dt : 0.0;
    return max_power < avail ? max_power : avail;
}

// ---------------------------------------------------------------------------
// Lifecycle (src/sim/lifecycle.cuh)
// ---------------------------------------------------------------------------
inline double co2_uptake_rate(double c) {                    // lifecycle.cuh:24-28
    if (c <= 0.0) return 0.0;
    return ASTRO_LIFE_CO2_UPTAKE_MAX * c / (c + ASTRO_LIFE_CO2_HALF_SATURATION);
}
inline double daughter_energy(double e) {                    // lifecycle.cuh:48-51
    const double half = 0.5 * (e - ASTRO_LIFE_DIVISION_ENERGY_COST);
    return half > 0.0 ? half : 0.0;
}
inline double corpse_energy(int d, double e) { return d == ASTRO_STORE_RETAIN ? e : 0.0; }  // :66-73
inline void division_axis(uint64_t did, double& ax, double& ay, double& az) {  // :87-99
    const uint64_t h = splitmix64(splitmix64(did));
    const double u = static_cast<double>(h >> 33) * (1.0 / 2147483648.0) - 1.0;
    const double phi = static_cast<double>(h & 0xFFFFFFFFull) * (ASTRO_TWO_PI / 4294967296.0);
    const double s = std::sqrt(u * u < 1.0 ? 1.0 - u * u : 0.0);
    ax = s * std::cos(phi);
    ay = s * std::sin(phi);
    az = u;
}

inline uint64_t fnv1a64(const void* data, size_t n, uint64_t h) {  // contracts/snapshot_v1.h:59-63
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) { h ^= p[i]; h *= 1099511628211ull; }
    return h;
}

// ---------------------------------------------------------------------------
// Field grid (src/fields/grid.cuh, src/fields/diffuse.cu)
// ---------------------------------------------------------------------------
struct Grid {
    bool   enabled = false;
    bool   square2d = false;          // reference geometry: n x n x 1
    int    nx = 0, ny = 0, nz = 0;
    double dx = 0, dy = 0, dz = 0;    // [m], double as the reference's Grid2D
    float  fdx = 0, fdy = 0, fdz = 0; // float copies, as FieldView carries them
    double D = 0, ambient = 0, scale = 1;
    int    bc = ASTRO_BC_NEUMANN;
    double robin_x = 0, robin_y = 0, robin_z = 0;
    std::vector<float> value, scratch;
    std::vector<uint64_t> deposit;

    size_t count() const { return static_cast<size_t>(nx) * ny * nz; }
    size_t at(int i, int j, int k) const { return (static_cast<size_t>(k) * ny + j) * nx + i; }
};

// grid_create (src/fields/diffuse.cu:91-119), extended to nz > 1.
void grid_init(Grid& g, astro_grid_dims dims, double w, double h, double d, double D,
               double ambient, double scale, int bc) {
    g.enabled = dims.nx > 0;
    if (!g.enabled) return;
    g.nx = dims.nx; g.ny = dims.ny; g.nz = dims.nz;
    g.square2d = (g.nz == 1 && g.nx == g.ny);
    g.dx = w / static_cast<double>(g.nx);
    g.dy = g.square2d ? g.dx : h / static_cast<double>(g.ny);
    g.dz = d / static_cast<double>(g.nz);
    g.fdx = static_cast<float>(g.dx);
    g.fdy = static_cast<float>(g.dy);
    g.fdz = static_cast<float>(g.dz);
    g.D = D; g.ambient = ambient; g.scale = scale; g.bc = bc;
    g.robin_x = g.dx * ASTRO_AIR_CONVECTION_H / ASTRO_WATER_CONDUCTIVITY;
    g.robin_y = g.dy * ASTRO_AIR_CONVECTION_H / ASTRO_WATER_CONDUCTIVITY;
    g.robin_z = g.dz * ASTRO_AIR_CONVECTION_H / ASTRO_WATER_CONDUCTIVITY;
    g.value.assign(g.count(), static_cast<float>(ambient));
    g.scratch.assign(g.count(), static_cast<float>(ambient));
    g.deposit.assign(g.count(), 0ull);
}

// substeps_for_dt (src/fields/grid.cuh:51-53); 3D uses the 7-point bound.
int substeps_for_dt(const Grid& g, double dt) {
    if (g.square2d) return static_cast<int>(dt / (g.dx * g.dx / (4.0 * g.D))) + 1;
    double inv = 1.0 / (g.dx * g.dx) + 1.0 / (g.dy * g.dy);
    if (g.nz > 1) inv += 1.0 / (g.dz * g.dz);
    return static_cast<int>(dt / (1.0 / (2.0 * g.D * inv))) + 1;
}

// Bilinear sample (src/fields/grid.cuh:109-125); trilinear when nz > 1.
float grid_sample(const Grid& g, float x, float y, float z) {
    const float half_x = 0.5f * g.nx * g.fdx;
    const float half_y = 0.5f * g.ny * g.fdy;
    const float gx = (x + half_x) / g.fdx - 0.5f;
    const float gy = (y + half_y) / g.fdy - 0.5f;
    const int x0 = static_cast<int>(std::floor(gx));
    const int y0 = static_cast<int>(std::floor(gy));
    const float tx = gx - static_cast<float>(x0);
    const float ty = gy - static_cast<float>(y0);
    const int x1 = x0 + 1, y1 = y0 + 1;
    auto cl = [](int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); };
    auto plane = [&](int k) -> float {
        auto at = [&](int i, int j) { return g.value[g.at(cl(i, g.nx), cl(j, g.ny), k)]; };
        return (1.0f - tx) * (1.0f - ty) * at(x0, y0) + tx * (1.0f - ty) * at(x1, y0) +
               (1.0f - tx) * ty * at(x0, y1) + tx * ty * at(x1, y1);
    };
    if (g.nz == 1) return plane(0);
    const float half_z = 0.5f * g.nz * g.fdz;
    const float gz = (z + half_z) / g.fdz - 0.5f;
    const int z0 = static_cast<int>(std::floor(gz));
    const float tz = gz - static_cast<float>(z0);
    const float b0 = plane(cl(z0, g.nz));
    const float b1 = plane(cl(z0 + 1, g.nz));
    return (1.0f - tz) * b0 + tz * b1;
}

// Nearest-cell index (src/fields/grid.cuh:134-141); a voxel when nz > 1.
size_t grid_nearest(const Grid& g, float x, float y, float z) {
    const float half_x = 0.5f * g.nx * g.fdx;
    const float half_y = 0.5f * g.ny * g.fdy;
    int i = static_cast<int>((x + half_x) / g.fdx);
    int j = static_cast<int>((y + half_y) / g.fdy);
    i = i < 0 ? 0 : (i >= g.nx ? g.nx - 1 : i);
    j = j < 0 ? 0 : (j >= g.ny ? g.ny - 1 : j);
    int k = 0;
    if (g.nz > 1) {
        const float half_z = 0.5f * g.nz * g.fdz;
        k = static_cast<int>((z + half_z) / g.fdz);
        k = k < 0 ? 0 : (k >= g.nz ? g.nz - 1 : k);
    }
    return g.at(i, j, k);
}

// flush_kernel (src/fields/diffuse.cu:51-60).
void grid_flush(Grid& g) {
    if (!g.enabled) return;
    const size_t count = g.count();
    PAR_FOR
    for (size_t i = 0; i < count; ++i) {
        const uint64_t raw = g.deposit[i];
        if (raw != 0ull) {
            g.value[i] += static_cast<float>(from_fixed(raw, g.scale));
            g.deposit[i] = 0ull;
        }
    }
}

// One explicit substep (diffuse_kernel, src/fields/diffuse.cu:19-44).
// Returns false when the coefficient is unstable (diffuse.cu:146, :172).
bool grid_substep(Grid& g, double dt_sub) {
    if (!g.enabled) return true;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const float amb = static_cast<float>(g.ambient);
    if (g.square2d) {
        const float coeff = static_cast<float>(g.D * dt_sub / (g.dx * g.dx));
        if (!(coeff <= 0.25f)) return false;
        const float robin = static_cast<float>(g.robin_x);
        const int n = nx;
        PAR_FOR
        for (int j = 0; j < n; ++j) {
            for (int i = 0; i < n; ++i) {
                const float c = g.value[j * n + i];
                auto ghost = [&]() -> float {
                    if (g.bc == ASTRO_BC_DIRICHLET) return amb;
                    if (g.bc == ASTRO_BC_ROBIN) return c - robin * (c - amb);
                    return c;
                };
                const float l = (i > 0) ? g.value[j * n + (i - 1)] : ghost();
                const float r = (i < n - 1) ? g.value[j * n + (i + 1)] : ghost();
                const float d = (j > 0) ? g.value[(j - 1) * n + i] : ghost();
                const float u = (j < n - 1) ? g.value[(j + 1) * n + i] : ghost();
                g.scratch[j * n + i] = c + coeff * (l + r + d + u - 4.0f * c);
            }
        }
    } else {
        const float cx = static_cast<float>(g.D * dt_sub / (g.dx * g.dx));
        const float cy = static_cast<float>(g.D * dt_sub / (g.dy * g.dy));
        const float cz = nz > 1 ? static_cast<float>(g.D * dt_sub / (g.dz * g.dz)) : 0.0f;
        if (!(static_cast<double>(cx) + cy + cz <= 0.5)) return false;
        const float rx = static_cast<float>(g.robin_x);
        const float ry = static_cast<float>(g.robin_y);
        const float rz = static_cast<float>(g.robin_z);
        PAR_FOR2
        for (int k = 0; k < nz; ++k) {
            for (int j = 0; j < ny; ++j) {
                for (int i = 0; i < nx; ++i) {
                    const size_t id = g.at(i, j, k);
                    const float c = g.value[id];
                    auto ghost = [&](float rob) -> float {
                        if (g.bc == ASTRO_BC_DIRICHLET) return amb;
                        if (g.bc == ASTRO_BC_ROBIN) return c - rob * (c - amb);
                        return c;
                    };
                    const float l = (i > 0) ? g.value[id - 1] : ghost(rx);
                    const float r = (i < nx - 1) ? g.value[id + 1] : ghost(rx);
                    const float d = (j > 0) ? g.value[id - nx] : ghost(ry);
                    const float u = (j < ny - 1) ? g.value[id + nx] : ghost(ry);
                    float acc = cx * ((l + r) - 2.0f * c) + cy * ((d + u) - 2.0f * c);
                    if (nz > 1) {
                        const size_t pl = static_cast<size_t>(nx) * ny;
                        const float b = (k > 0) ? g.value[id - pl] : ghost(rz);
                        const float f = (k < nz - 1) ? g.value[id + pl] : ghost(rz);
                        acc = acc + cz * ((b + f) - 2.0f * c);
                    }
                    g.scratch[id] = c + acc;
                }
            }
        }
    }
    std::swap(g.value, g.scratch);
    return true;
}

// grid_diffuse (src/fields/diffuse.cu:159-186).
bool grid_diffuse(Grid& g, double dt) {
    if (!g.enabled) return true;
    const int steps = substeps_for_dt(g, dt);
    if (steps <= 0) return true;
    const double dt_sub = dt / steps;
    for (int s = 0; s < steps; ++s)
        if (!grid_substep(g, dt_sub)) return false;
    return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// The world
// ---------------------------------------------------------------------------
struct oracle_world {
    astro_world_desc desc{};
    // cell SoA, contract order (contracts/cell_store_v1.h:42-74)
    std::vector<uint64_t> id, rng_state;
    std::vector<uint32_t> flags;
    std::vector<uint8_t>  death_cause;
    std::vector<double>   x, y, z, vx, vy, vz, energy, biomass, co2_held;
    std::vector<float>    temp_cell, emit_power, dir_x, dir_y, dir_z, age_s, taxis_memory,
                          run_timer, t_local, irradiance, co2_local, n2_local;
    int32_t  capacity = 0, count = 0;
    uint64_t next_id = 1, tick = 0;
    double   sim_time = 0.0, physics_rate = 1.0, biology_rate = 1.0;
    int64_t  divisions_total = 0, deaths_total = 0;
    int64_t  divisions_reported = 0, deaths_reported = 0;
    int32_t  dead_slots_prev = 0;
    int64_t  cell_steps = 0;

    Grid f[ASTRO_FIELD_COUNT];

    // spatial hash (src/sim/hash.cuh:25-48)
    int32_t hnx = 0, hny = 0, hnz = 0, nb = 0;
    double  hcell = 0.0;
    float   h_inv = 0.f, h_ox = 0.f, h_oy = 0.f, h_oz = 0.f;
    std::vector<uint32_t> vals_sorted;
    std::vector<int32_t>  bstart, bend;
    bool    hash_valid = false;

    std::vector<double> fx, fy, fz;
    std::vector<uint64_t> co2_demand;

    // TaumoebaStore (src/sim/predation.cuh:85-126)
    struct Tau {
        std::vector<uint64_t> id, rng_state;
        std::vector<uint32_t> flags, generation;
        std::vector<double> x, y, z, vx, vy, vz, biomass, prey_biomass;
        std::vector<float> dir_x, dir_y, dir_z, digest_timer, tolerance, density_ema, run_timer;
        std::vector<int32_t> target;
        int32_t count = 0, capacity = 0;
        uint64_t next_id = 1;
    } tau;
    std::vector<uint64_t> claim;   // World::d_predator_claim, per cell slot
    bool flash_active = false;     // World::flash_active
    uint64_t flash_accum[4] = {0, 0, 0, 0};
    // Test-only: round FP32-layout positions and velocities to float where
    // the device layout stores them (oracle_world_set_float_store).
    bool float_store = false;

    std::string err;
};

namespace {

// A position or velocity as the world stores it.
inline double stored(const oracle_world* w, double v) {
    return (w->float_store && w->desc.precision == ASTRO_FP32_HOT) ? static_cast<double>(static_cast<float>(v)) : v;
}

int fail(oracle_world* w, int32_t st, const char* what) {
    if (w) w->err = what;
    return st;
}

void resize_cells(oracle_world* w, int32_t cap) {
    w->id.assign(cap, 0); w->rng_state.assign(cap, 0); w->flags.assign(cap, 0);
    w->death_cause.assign(cap, 0);
    for (auto* v : {&w->x, &w->y, &w->z, &w->vx, &w->vy, &w->vz, &w->energy, &w->biomass,
                    &w->co2_held})
        v->assign(cap, 0.0);
    for (auto* v : {&w->temp_cell, &w->emit_power, &w->dir_x, &w->dir_y, &w->dir_z, &w->age_s,
                    &w->taxis_memory, &w->run_timer, &w->t_local, &w->irradiance, &w->co2_local,
                    &w->n2_local})
        v->assign(cap, 0.0f);
    w->fx.assign(cap, 0.0); w->fy.assign(cap, 0.0); w->fz.assign(cap, 0.0);
    w->vals_sorted.assign(cap, 0);
}

inline V3 pos_of(const oracle_world* w, int i) { return {w->x[i], w->y[i], w->z[i]}; }

// hash_coords (src/sim/hash.cuh:63-68)
inline void hash_coords(const oracle_world* w, double px, double py, double pz,
                        int32_t& gx, int32_t& gy, int32_t& gz) {
    auto cl = [](int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); };
    gx = cl(static_cast<int32_t>((px - w->h_ox) * w->h_inv), 0, w->hnx - 1);
    gy = cl(static_cast<int32_t>((py - w->h_oy) * w->h_inv), 0, w->hny - 1);
    gz = cl(static_cast<int32_t>((pz - w->h_oz) * w->h_inv), 0, w->hnz - 1);
}
inline uint32_t hash_index(const oracle_world* w, int32_t gx, int32_t gy, int32_t gz) {
    return static_cast<uint32_t>((gz * w->hny + gy) * w->hnx + gx);
}

// ASTRO_FOR_EACH_NEIGHBOUR (src/sim/hash.cuh:88-111): z, then y, then x,
// then sorted position within the bucket.
template <class F>
void for_each_neighbour(const oracle_world* w, double px, double py, double pz, F&& fn) {
    int32_t gx, gy, gz;
    hash_coords(w, px, py, pz, gx, gy, gz);
    for (int32_t dz = -1; dz <= 1; ++dz) {
        const int32_t cz = gz + dz;
        if (cz < 0 || cz >= w->hnz) continue;
        for (int32_t dy = -1; dy <= 1; ++dy) {
            const int32_t cy = gy + dy;
            if (cy < 0 || cy >= w->hny) continue;
            for (int32_t dx = -1; dx <= 1; ++dx) {
                const int32_t cx = gx + dx;
                if (cx < 0 || cx >= w->hnx) continue;
                const uint32_t b = hash_index(w, cx, cy, cz);
                for (int32_t k = w->bstart[b]; k < w->bend[b]; ++k)
                    fn(static_cast<int32_t>(w->vals_sorted[k]));
            }
        }
    }
}

// hash_build (src/sim/hash.cu:106-133): keys, stable sort, ranges.
void hash_build(oracle_world* w) {
    const int32_t n = w->count;
    std::fill(w->bstart.begin(), w->bstart.end(), 0);
    std::fill(w->bend.begin(), w->bend.end(), 0);
    w->hash_valid = true;
    if (n <= 0) return;
    std::vector<uint32_t> keys(n);
    PAR_FOR
    for (int32_t i = 0; i < n; ++i) {
        if (w->flags[i] & ASTRO_FLAG_OCCUPIED) {
            int32_t gx, gy, gz;
            hash_coords(w, w->x[i], w->y[i], w->z[i], gx, gy, gz);
            keys[i] = hash_index(w, gx, gy, gz);
        } else {
            keys[i] = static_cast<uint32_t>(w->nb);
        }
    }
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    STABLE_SORT(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    for (int32_t k = 0; k < n; ++k) w->vals_sorted[k] = order[k];
    for (int32_t k = 0; k < n; ++k) {                          // range_kernel, hash.cu:29-38
        const uint32_t key = keys[order[k]];
        if (key >= static_cast<uint32_t>(w->nb)) continue;
        if (k == 0 || keys[order[k - 1]] != key) w->bstart[key] = k;
        if (k == n - 1 || keys[order[k + 1]] != key) w->bend[key] = k + 1;
    }
}

// ---------------------------------------------------------------------------
// Stages
// ---------------------------------------------------------------------------

// thermal_step (src/sim/thermal.cu:75-103) with thermal_kernel (:19-71).
int thermal_step(oracle_world* w, double dt) {
    Grid& T = w->f[ASTRO_FIELD_TEMPERATURE];
    const int32_t n = w->count;
    const int substeps = substeps_for_dt(T, dt);
    const double dt_sub = dt / substeps;
    const double cap = ASTRO_WATER_DENSITY * ASTRO_WATER_SPECIFIC_HEAT * T.dx * T.dy * T.dz;
    const double g = ASTRO_CONDUCTION_COEFF;
    for (int s = 0; s < substeps; ++s) {
        if (n > 0 && w->desc.thermal_enabled) {
            PAR_FOR
            for (int32_t i = 0; i < n; ++i) {
                const uint32_t fl = w->flags[i];
                if (!(fl & ASTRO_FLAG_OCCUPIED)) continue;
                const float fxp = static_cast<float>(w->x[i]);
                const float fyp = static_cast<float>(w->y[i]);
                const float fzp = static_cast<float>(w->z[i]);
                const size_t idx = grid_nearest(T, fxp, fyp, fzp);
                const float t_grid = T.value[idx];
                bool awake = (fl & ASTRO_FLAG_AWAKE) != 0u;
                if (!awake && t_grid >= ASTRO_CELL_TEMP_SETPOINT) {
                    awake = true;
                    w->flags[i] = fl | ASTRO_FLAG_AWAKE;
                }
                const double t_cell = awake ? ASTRO_CELL_TEMP_SETPOINT : static_cast<double>(t_grid);
                w->temp_cell[i] = static_cast<float>(t_cell);
                if (!awake || !(fl & ASTRO_FLAG_ALIVE)) continue;
                const double e = lumped_exchange_energy(t_cell, static_cast<double>(t_grid), g, cap, dt_sub);
                double energy = w->energy[i] - e;
                double moved = e;
                if (energy < 0.0) { moved = w->energy[i]; energy = 0.0; }
                else if (energy > ASTRO_CELL_ENERGY_MAX) {
                    moved = w->energy[i] - ASTRO_CELL_ENERGY_MAX;
                    energy = ASTRO_CELL_ENERGY_MAX;
                }
                w->energy[i] = energy;
                FIXED_ADD(T.deposit[idx], static_cast<uint64_t>(to_fixed(moved / cap, T.scale)));
                if (energy <= 0.0) {
                    w->flags[i] = w->flags[i] & ~ASTRO_FLAG_ALIVE;
                    w->death_cause[i] = ASTRO_DEATH_STARVED;
                    w->emit_power[i] = 0.0f;
                }
            }
            grid_flush(T);
        }
        if (!grid_substep(T, dt_sub)) return fail(w, ASTRO_INVALID_ARGUMENT, "diffusion coefficient unstable");
    }
    return ASTRO_OK;
}

// emission_step, directional path (src/sim/emission.cu:119-177).
int emission_step(oracle_world* w, double dt) {
    Grid& I = w->f[ASTRO_FIELD_IRRADIANCE];
    const astro_world_desc& d = w->desc;
    const float source = d.light_enabled ? d.light_irradiance : 0.0f;
    const float ambient = d.ambient_irradiance;
    if (source <= 0.0f && ambient <= 0.0f) return ASTRO_OK;
    const V3 dir = normalize(V3{static_cast<double>(d.light_dir_x), static_cast<double>(d.light_dir_y), 0.0});
    std::fill(I.deposit.begin(), I.deposit.end(), 0ull);       // clear_kernel
    const int32_t n = w->count;
    if (n > 0 && source > 0.0f) {                                // occlusion_stamp_kernel
        PAR_FOR
        for (int32_t i = 0; i < n; ++i) {
            const uint32_t fl = w->flags[i];
            if (!(fl & ASTRO_FLAG_OCCUPIED)) continue;
            const double area = (fl & ASTRO_FLAG_ALIVE) ? ASTRO_CELL_CROSS_SECTION
                                : ASTRO_CELL_CROSS_SECTION * (1.0 - ASTRO_CELL_ALBEDO_DEAD);
            const size_t idx = grid_nearest(I, static_cast<float>(w->x[i]), static_cast<float>(w->y[i]),
                                            static_cast<float>(w->z[i]));
            FIXED_ADD(I.deposit[idx], static_cast<uint64_t>(to_fixed(area, I.scale)));
        }
    }
    const int axis = (std::fabs(dir.x) >= std::fabs(dir.y)) ? 0 : 1;
    const int sign = (axis == 0) ? (dir.x >= 0.0 ? 1 : -1) : (dir.y >= 0.0 ? 1 : -1);
    const double face = (axis == 0 ? I.dy : I.dx) * I.dz;
    // sweep_kernel (emission.cu:37-53): one sequential march per line.
    const int len = axis == 0 ? I.nx : I.ny;
    const int other = axis == 0 ? I.ny : I.nx;
    PAR_FOR2
    for (int k = 0; k < I.nz; ++k) {
        for (int line = 0; line < other; ++line) {
            double t = 1.0;
            for (int s = 0; s < len; ++s) {
                const int step = sign > 0 ? s : (len - 1 - s);
                const size_t idx = axis == 0 ? I.at(step, line, k) : I.at(line, step, k);
                I.value[idx] = static_cast<float>(source * t) + ambient;
                const double area = from_fixed(I.deposit[idx], I.scale);
                t *= column_transmittance(area, face);
            }
        }
    }
    // feed_kernel (emission.cu:81-115)
    if (n > 0) {
        const V3 ld = dir;
        PAR_FOR
        for (int32_t i = 0; i < n; ++i) {
            const uint32_t fl = w->flags[i];
            if (!(fl & ASTRO_FLAG_OCCUPIED) || !(fl & ASTRO_FLAG_ALIVE)) continue;
            const V3 p = pos_of(w, i);
            double directional = static_cast<double>(grid_sample(I, static_cast<float>(p.x),
                static_cast<float>(p.y), static_cast<float>(p.z))) - ambient;
            if (directional < 0.0) directional = 0.0;
            if (d.occlusion_exact && w->hash_valid) {
                double open = 1.0;
                for_each_neighbour(w, p.x, p.y, p.z, [&](int32_t j) {
                    if (j != i && (w->flags[j] & ASTRO_FLAG_OCCUPIED) && (w->flags[j] & ASTRO_FLAG_ALIVE))
                        open *= (1.0 - shadow_fraction(p, pos_of(w, j), ld));
                });
                directional *= open;
            }
            w->irradiance[i] = static_cast<float>(directional);
            const double gained = absorbed_power(directional, static_cast<double>(ambient)) * dt;
            double e = w->energy[i] + gained;
            if (e > ASTRO_CELL_ENERGY_MAX) e = ASTRO_CELL_ENERGY_MAX;
            w->energy[i] = e;
        }
    }
    return ASTRO_OK;
}

// taxis_kernel (src/sim/taxis.cu:13-87).
void taxis_step(oracle_world* w, double dt) {
    const astro_world_desc& d = w->desc;
    PAR_FOR
    for (int32_t i = 0; i < w->count; ++i) {
        const uint32_t fl = w->flags[i];
        if (!(fl & ASTRO_FLAG_OCCUPIED) || !(fl & ASTRO_FLAG_ALIVE)) continue;
        if (!(fl & ASTRO_FLAG_AWAKE)) { w->emit_power[i] = 0.0f; continue; }
        const double irr = static_cast<double>(w->irradiance[i]);
        const double co2 = static_cast<double>(w->co2_local[i]);
        const double charge = w->energy[i] / ASTRO_CELL_ENERGY_MAX;
        const int state = taxis_select_state(charge, irr, co2);
        if (state == IDLE) {
            w->emit_power[i] = 0.0f;
            w->run_timer[i] = 0.0f;
            w->taxis_memory[i] = 0.0f;
            continue;
        }
        const double signal = state == FEED ? irr : co2;
        const double delta = signal - static_cast<double>(w->taxis_memory[i]);
        w->taxis_memory[i] = taxis_ema_update(w->taxis_memory[i], signal, dt);
        V3 heading = -V3{static_cast<double>(w->dir_x[i]), static_cast<double>(w->dir_y[i]),
                         static_cast<double>(w->dir_z[i])};
        double run = static_cast<double>(w->run_timer[i]) + dt;
        if (delta <= 0.0 || run >= ASTRO_TAXIS_RUN_MAX) {
            if (d.rng_kind == ASTRO_RNG_PHILOX) {
                Rng r = philox_stream(d.seed, ASTRO_RNG_STAGE_TAXIS, w->tick, w->id[i]);
                heading = taxis_tumble(heading, r);
            } else {
                Rng r = cell_rng(w->rng_state[i], w->id[i]);
                heading = taxis_tumble(heading, r);
                w->rng_state[i] = r.state;
            }
            run = 0.0;
        }
        w->run_timer[i] = static_cast<float>(run);
        const V3 axis = -normalize(heading);
        w->dir_x[i] = static_cast<float>(axis.x);
        w->dir_y[i] = static_cast<float>(axis.y);
        w->dir_z[i] = static_cast<float>(axis.z);
        const double p = taxis_emit_power(state, w->energy[i], dt, w->desc.petrova_max_power);
        w->emit_power[i] = static_cast<float>(p);
        w->energy[i] -= p * dt;
    }
}

// field_sample_kernel (src/sim/integrator.cu:20-30).
void field_sample(oracle_world* w) {
    PAR_FOR
    for (int32_t i = 0; i < w->count; ++i) {
        if (!(w->flags[i] & ASTRO_FLAG_OCCUPIED)) continue;
        const float x = static_cast<float>(w->x[i]);
        const float y = static_cast<float>(w->y[i]);
        const float z = static_cast<float>(w->z[i]);
        w->t_local[i] = grid_sample(w->f[ASTRO_FIELD_TEMPERATURE], x, y, z);
        w->co2_local[i] = w->f[ASTRO_FIELD_CO2].enabled ? grid_sample(w->f[ASTRO_FIELD_CO2], x, y, z) : 0.0f;
        w->n2_local[i] = w->f[ASTRO_FIELD_N2].enabled ? grid_sample(w->f[ASTRO_FIELD_N2], x, y, z) : 0.0f;
    }
}

// forces_kernel (src/sim/integrator.cu:40-76).
void forces(oracle_world* w) {
    const double k = ASTRO_CONTACT_STIFFNESS / w->physics_rate;
    PAR_FOR
    for (int32_t i = 0; i < w->count; ++i) {
        const uint32_t fl = w->flags[i];
        if (!(fl & ASTRO_FLAG_OCCUPIED)) { w->fx[i] = w->fy[i] = w->fz[i] = 0.0; continue; }
        const double mass = cell_mass(w->biomass[i], w->energy[i]);
        const bool alive = (fl & ASTRO_FLAG_ALIVE) != 0u;
        const double emit = alive ? static_cast<double>(w->emit_power[i]) : 0.0;
        const V3 dir{w->dir_x[i], w->dir_y[i], w->dir_z[i]};
        V3 f{0.0, 0.0, 0.0};                                     // cell_force, integrator.cuh:179-188
        const double wgt = buoyant_weight(mass);
        if (w->desc.gravity_axis == ASTRO_GRAVITY_Y) f.y += wgt; else f.z += wgt;
        const double thrust = emit / ASTRO_C_LIGHT;
        if (thrust != 0.0) f = f - dir * thrust;
        if (w->desc.contact_enabled && w->hash_valid) {
            const V3 p = pos_of(w, i);
            for_each_neighbour(w, p.x, p.y, p.z, [&](int32_t j) {
                if (j != i && (w->flags[j] & ASTRO_FLAG_OCCUPIED)) f = f + contact_force(p, pos_of(w, j), k);
            });
        }
        w->fx[i] = f.x; w->fy[i] = f.y; w->fz[i] = f.z;
    }
}

// integrate_kernel (src/sim/integrator.cu:79-150).
void integrate(oracle_world* w, double dt) {
    const astro_world_desc& d = w->desc;
    PAR_FOR
    for (int32_t i = 0; i < w->count; ++i) {
        const uint32_t fl = w->flags[i];
        if (!(fl & ASTRO_FLAG_OCCUPIED)) continue;
        const double t_local = static_cast<double>(w->t_local[i]);
        const double mass = cell_mass(w->biomass[i], w->energy[i]);
        const bool awake = (fl & ASTRO_FLAG_AWAKE) != 0u;
        const double t_drag = awake ? ASTRO_CELL_TEMP_SETPOINT : t_local;
        double gamma = drag_coefficient(t_drag);
        const V3 force{w->fx[i], w->fy[i], w->fz[i]};
        V3 pos = pos_of(w, i);
        V3 vel{w->vx[i], w->vy[i], w->vz[i]};
        if (fl & ASTRO_FLAG_STUCK) gamma *= ASTRO_WALL_STUCK_DRAG_MULT;
        const bool philox = d.rng_kind == ASTRO_RNG_PHILOX;
        Rng r = philox ? philox_stream(d.seed, ASTRO_RNG_STAGE_INTEGRATE, w->tick, w->id[i])
                       : cell_rng(w->rng_state[i], w->id[i]);
        const Ou c = ou_coefficients(mass, gamma, t_drag, dt);   // integrate_cell, integrator.cuh:158-169
        double z1, z2, z3, z4, z5, z6;
        gaussian_pair(r, z1, z2);
        gaussian_pair(r, z3, z4);
        gaussian_pair(r, z5, z6);
        const bool noise = d.thermal_noise != 0;
        ou_advance_axis(pos.x, vel.x, force.x, gamma, c, dt, noise, z1, z2);
        ou_advance_axis(pos.y, vel.y, force.y, gamma, c, dt, noise, z3, z4);
        ou_advance_axis(pos.z, vel.z, force.z, gamma, c, dt, noise, z5, z6);
        if (!philox) w->rng_state[i] = r.state;
        const double a = ASTRO_CELL_RADIUS;
        bool kept = apply_boundary_axis(pos.x, vel.x, 0.5 * d.chamber_w, a, d.boundary_x);
        kept = apply_boundary_axis(pos.y, vel.y, 0.5 * d.chamber_h, a, d.boundary_y) && kept;
        const double z_before = pos.z;
        apply_boundary_axis(pos.z, vel.z, 0.5 * d.chamber_d, a, ASTRO_BOUNDARY_REFLECTING);
        const bool touched = (pos.z != z_before);
        uint32_t nf = fl;
        if (d.adhesion_enabled) {
            if (touched && !(fl & ASTRO_FLAG_STUCK)) {
                // PCG32: re-derived from the state just stored (integrator.cu:
                // 124-127); Philox: the stage stream's next draws
                bool sticks;
                if (philox) {
                    sticks = uniform01d(r) < ASTRO_WALL_STICKINESS;
                } else {
                    Rng ar = cell_rng(w->rng_state[i], w->id[i]);
                    sticks = uniform01d(ar) < ASTRO_WALL_STICKINESS;
                    w->rng_state[i] = ar.state;
                }
                if (sticks) nf |= ASTRO_FLAG_STUCK;
            } else if (fl & ASTRO_FLAG_STUCK) {
                const double away = (pos.z < 0.0) ? force.z : -force.z;
                if (away > ASTRO_WALL_RELEASE_FORCE) nf &= ~ASTRO_FLAG_STUCK;
            }
        }
        w->flags[i] = nf;
        if (!kept) {
            w->flags[i] = nf & ~ASTRO_FLAG_ALIVE;
            w->death_cause[i] = ASTRO_DEATH_CULLED;
            w->emit_power[i] = 0.0f;
        }
        w->x[i] = stored(w, pos.x); w->y[i] = stored(w, pos.y); w->z[i] = stored(w, pos.z);
        w->vx[i] = stored(w, vel.x); w->vy[i] = stored(w, vel.y); w->vz[i] = stored(w, vel.z);
        w->age_s[i] += static_cast<float>(dt);
    }
}

// cell_store_compact (src/sim/cell_store.cu:414-456): stable, out of place.
template <class T>
void compact_array(std::vector<T>& a, const std::vector<int32_t>& src, int32_t live) {
    std::vector<T> tmp(live);
    for (int32_t k = 0; k < live; ++k) tmp[k] = a[src[k]];
    std::copy(tmp.begin(), tmp.end(), a.begin());
}

void compact(oracle_world* w) {
    const int32_t cnt = w->count;
    if (cnt <= 0) return;
    std::vector<int32_t> src;
    src.reserve(cnt);
    for (int32_t i = 0; i < cnt; ++i) {
        const uint32_t f = w->flags[i];
        if ((f & ASTRO_FLAG_OCCUPIED) && (f & ASTRO_FLAG_ALIVE)) src.push_back(i);
    }
    const int32_t live = static_cast<int32_t>(src.size());
    if (live >= cnt) return;
    compact_array(w->id, src, live);        compact_array(w->flags, src, live);
    compact_array(w->death_cause, src, live); compact_array(w->x, src, live);
    compact_array(w->y, src, live);         compact_array(w->z, src, live);
    compact_array(w->vx, src, live);        compact_array(w->vy, src, live);
    compact_array(w->vz, src, live);        compact_array(w->energy, src, live);
    compact_array(w->temp_cell, src, live); compact_array(w->emit_power, src, live);
    compact_array(w->dir_x, src, live);     compact_array(w->dir_y, src, live);
    compact_array(w->dir_z, src, live);     compact_array(w->biomass, src, live);
    compact_array(w->co2_held, src, live);  compact_array(w->age_s, src, live);
    compact_array(w->rng_state, src, live); compact_array(w->taxis_memory, src, live);
    compact_array(w->run_timer, src, live); compact_array(w->t_local, src, live);
    compact_array(w->irradiance, src, live); compact_array(w->co2_local, src, live);
    compact_array(w->n2_local, src, live);
    w->count = live;
}

// lifecycle_step (src/sim/lifecycle.cu:208-294).
void lifecycle_step(oracle_world* w, double dt) {
    const int32_t n = w->count;
    if (n <= 0) return;
    const double dt_bio = dt * w->biology_rate;
    Grid& C = w->f[ASTRO_FIELD_CO2];
    const astro_world_desc& d = w->desc;
    const double quota = d.co2_mass_per_division;
    if (C.enabled) {
        const double gv = C.dx * C.dy * C.dz;
        std::fill(w->co2_demand.begin(), w->co2_demand.end(), 0ull);     // clear_u64_kernel
        auto want_of = [&](int32_t i, size_t& idx) {                      // co2_demand, :27-33
            idx = grid_nearest(C, static_cast<float>(w->x[i]), static_cast<float>(w->y[i]),
                               static_cast<float>(w->z[i]));
            return co2_uptake_rate(static_cast<double>(C.value[idx])) * dt_bio;
        };
        PAR_FOR
        for (int32_t i = 0; i < n; ++i) {                                  // demand_kernel
            const uint32_t f = w->flags[i];
            if (!(f & ASTRO_FLAG_OCCUPIED) || !(f & ASTRO_FLAG_ALIVE)) continue;
            size_t idx;
            const double want = want_of(i, idx);
            if (want <= 0.0) continue;
            FIXED_ADD(w->co2_demand[idx], static_cast<uint64_t>(to_fixed(want / gv, C.scale)));
        }
        PAR_FOR
        for (int32_t i = 0; i < n; ++i) {                                  // uptake_kernel
            const uint32_t f = w->flags[i];
            if (!(f & ASTRO_FLAG_OCCUPIED) || !(f & ASTRO_FLAG_ALIVE)) continue;
            size_t idx;
            const double want = want_of(i, idx);
            if (want <= 0.0) continue;
            const double asked = from_fixed(w->co2_demand[idx], C.scale);
            const double available = static_cast<double>(C.value[idx]);
            double taken = want;
            if (asked > available && asked > 0.0) taken = want * (available / asked);
            if (taken <= 0.0) continue;
            w->co2_held[i] += taken;
            FIXED_ADD(C.deposit[idx], static_cast<uint64_t>(to_fixed(-taken / gv, C.scale)));
        }
        grid_flush(C);
    }
    PAR_FOR
    for (int32_t i = 0; i < n; ++i) {                                      // death_kernel
        const uint32_t f = w->flags[i];
        if (!(f & ASTRO_FLAG_OCCUPIED) || !(f & ASTRO_FLAG_ALIVE)) continue;
        if (!(static_cast<double>(w->t_local[i]) > ASTRO_CELL_LETHAL_TEMP)) continue;
        w->flags[i] = f & ~ASTRO_FLAG_ALIVE;
        w->death_cause[i] = ASTRO_DEATH_OVERHEATED;
        w->emit_power[i] = 0.0f;
        w->energy[i] = corpse_energy(d.store_disposition, w->energy[i]);
    }
    std::vector<int32_t> divides(n, 0);                                    // mark_kernel
    int32_t births = 0, deads = 0;
    for (int32_t i = 0; i < n; ++i) {
        const uint32_t f = w->flags[i];
        const bool occ = (f & ASTRO_FLAG_OCCUPIED) != 0u, alive = (f & ASTRO_FLAG_ALIVE) != 0u;
        divides[i] = (occ && alive && w->co2_held[i] >= quota) ? 1 : 0;
        births += divides[i];
        if (occ && !alive) ++deads;
    }
    const int32_t new_deaths = deads - w->dead_slots_prev;
    if (new_deaths > 0) w->deaths_total += new_deaths;
    if (births > 0) {
        std::vector<int32_t> prefix(n, 0);
        std::exclusive_scan(divides.begin(), divides.end(), prefix.begin(), 0);
        const int32_t base = n;
        if (base + births > w->capacity) births = w->capacity - base;
        if (births > 0) {
            const double a = d.division_offset;
            const double hx = 0.5 * d.chamber_w - ASTRO_CELL_RADIUS;
            const double hy = 0.5 * d.chamber_h - ASTRO_CELL_RADIUS;
            const double hz = 0.5 * d.chamber_d - ASTRO_CELL_RADIUS;
            auto clampv = [](double v, double h) { return v < -h ? -h : (v > h ? h : v); };
            PAR_FOR
            for (int32_t i = 0; i < n; ++i) {                              // divide_kernel, :140-204
                const uint32_t f = w->flags[i];
                if (!(f & ASTRO_FLAG_OCCUPIED) || !(f & ASTRO_FLAG_ALIVE)) continue;
                if (!(w->co2_held[i] >= quota)) continue;
                const int32_t slot = base + prefix[i];
                if (slot >= w->capacity) continue;
                const uint64_t did = w->next_id + static_cast<uint64_t>(prefix[i]);
                const double half = daughter_energy(w->energy[i]);
                w->energy[i] = half;
                w->energy[slot] = half;
                w->co2_held[i] = 0.0;
                w->co2_held[slot] = 0.0;
                w->biomass[slot] = ASTRO_CELL_MASS_DRY;
                double ax, ay, az;
                division_axis(did, ax, ay, az);
                w->x[slot] = w->x[i] + a * ax;  w->x[i] -= a * ax;
                w->y[slot] = w->y[i] + a * ay;  w->y[i] -= a * ay;
                w->z[slot] = w->z[i] + a * az;  w->z[i] -= a * az;
                w->x[slot] = clampv(w->x[slot], hx); w->y[slot] = clampv(w->y[slot], hy);
                w->z[slot] = clampv(w->z[slot], hz);
                w->x[i] = clampv(w->x[i], hx); w->y[i] = clampv(w->y[i], hy);
                w->z[i] = clampv(w->z[i], hz);
                for (size_t j : {static_cast<size_t>(slot), static_cast<size_t>(i)}) {
                    w->x[j] = stored(w, w->x[j]); w->y[j] = stored(w, w->y[j]); w->z[j] = stored(w, w->z[j]);
                }
                w->vx[slot] = w->vx[i]; w->vy[slot] = w->vy[i]; w->vz[slot] = w->vz[i];
                w->id[slot] = did;
                w->flags[slot] = w->flags[i];
                w->death_cause[slot] = ASTRO_DEATH_NONE;
                w->temp_cell[slot] = w->temp_cell[i];
                w->emit_power[slot] = 0.0f;
                w->dir_x[slot] = w->dir_x[i]; w->dir_y[slot] = w->dir_y[i]; w->dir_z[slot] = w->dir_z[i];
                w->age_s[slot] = 0.0f;
                // a counter stream is keyed by the daughter's own id
                w->rng_state[slot] = d.rng_kind == ASTRO_RNG_PHILOX ? 0ull : pcg_split(w->rng_state[i], did);
                w->taxis_memory[slot] = 0.0f;
                w->run_timer[slot] = 0.0f;
                w->t_local[slot] = w->t_local[i];
                w->irradiance[slot] = 0.0f;
                w->co2_local[slot] = w->co2_local[i];
                w->n2_local[slot] = w->n2_local[i];
            }
            w->divisions_total += births;
            w->next_id += static_cast<uint64_t>(births);
            w->count += births;
        }
    }
    bool compacted = false;
    if (d.compaction_enabled && deads > 0) {
        compact(w);
        compacted = true;
    }
    w->dead_slots_prev = compacted ? 0 : deads;
}

int slow_fields(oracle_world* w, double dt) {                    // step.cu:168-171
    grid_flush(w->f[ASTRO_FIELD_CO2]);
    grid_flush(w->f[ASTRO_FIELD_N2]);
    if (!grid_diffuse(w->f[ASTRO_FIELD_CO2], dt)) return fail(w, ASTRO_INVALID_ARGUMENT, "diffusion coefficient unstable");
    if (!grid_diffuse(w->f[ASTRO_FIELD_N2], dt)) return fail(w, ASTRO_INVALID_ARGUMENT, "diffusion coefficient unstable");
    return ASTRO_OK;
}

// ---------------------------------------------------------------------------
// Taumoeba (src/sim/predation.cu, src/sim/predation.cuh)
// ---------------------------------------------------------------------------
void resize_tau(oracle_world::Tau& t, int32_t cap) {
    t.capacity = cap;
    t.id.assign(cap, 0); t.rng_state.assign(cap, 0); t.flags.assign(cap, 0); t.generation.assign(cap, 0);
    for (auto* v : {&t.x, &t.y, &t.z, &t.vx, &t.vy, &t.vz, &t.biomass, &t.prey_biomass}) v->assign(cap, 0.0);
    for (auto* v : {&t.dir_x, &t.dir_y, &t.dir_z, &t.digest_timer, &t.tolerance, &t.density_ema, &t.run_timer})
        v->assign(cap, 0.0f);
    t.target.assign(cap, -1);
}

inline bool live_flags(uint32_t f) { return (f & ASTRO_FLAG_OCCUPIED) && (f & ASTRO_FLAG_ALIVE); }

// The predator's pure helpers (src/sim/predation.cuh:24-80).
inline double tau_drag(double t) { return 6.0 * ASTRO_PI * water_viscosity(t) * ASTRO_TAU_RADIUS; }
inline bool tau_overlaps_cell(V3 tau, V3 cell) {
    const double r = ASTRO_TAU_RADIUS + ASTRO_CELL_RADIUS;
    const V3 d{tau.x - cell.x, tau.y - cell.y, tau.z - cell.z};
    return dot(d, d) < r * r;
}
inline bool tau_should_tumble(float signal, float& ema, float run_timer, double dt) {
    const double alpha = -std::expm1(-dt / ASTRO_TAXIS_MEMORY_TIME);
    const double delta = static_cast<double>(signal) - static_cast<double>(ema);
    ema = static_cast<float>(ema + alpha * delta);
    return !(delta > 0.0 && run_timer < ASTRO_TAXIS_RUN_MAX);
}
inline double tau_n2_hazard(double n2_local, double tolerance) {
    const double threshold = ASTRO_TAU_N2_LETHAL_CONC * (1.0 + tolerance * ASTRO_TAU_N2_TOLERANCE_K);
    const double h = n2_local - threshold;
    return h > 0.0 ? h : 0.0;
}
inline double tau_n2_death_prob(double hazard, double dt_bio) {
    return -std::expm1(-hazard * ASTRO_TAU_N2_HAZARD_RATE * dt_bio);
}
inline float tau_daughter_tolerance(float parent, double mutation) {
    const double v = static_cast<double>(parent) + ASTRO_TAU_MUTATION_SIGMA * mutation;
    return static_cast<float>(v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v));
}

// gaussian3 (src/core/rng.cuh:111-117) then normalize (src/core/vec.cuh:32-35).
inline V3 random_heading(Rng& r) {
    double a, b, c, d;
    gaussian_pair(r, a, b);
    gaussian_pair(r, c, d);
    return normalize(V3{a, b, c});
}

// spawn_kernel (src/sim/predation.cu:33-76).
void tau_spawn(oracle_world* w, int32_t count, uint64_t seed) {
    oracle_world::Tau& t = w->tau;
    const astro_world_desc& d = w->desc;
    for (int32_t i = 0; i < count; ++i) {
        const int32_t s = t.count + i;
        const uint64_t id = t.next_id + static_cast<uint64_t>(i);
        Rng r = cell_rng(cell_rng_init(splitmix64(seed), id), id);
        const double a = ASTRO_TAU_RADIUS;
        t.x[s] = uniform_range(r, -0.5 * d.chamber_w + a, 0.5 * d.chamber_w - a);
        t.y[s] = uniform_range(r, -0.5 * d.chamber_h + a, 0.5 * d.chamber_h - a);
        t.z[s] = uniform_range(r, -0.5 * d.chamber_d + a, 0.5 * d.chamber_d - a);
        t.vx[s] = 0.0; t.vy[s] = 0.0; t.vz[s] = 0.0;
        const V3 dir = random_heading(r);
        t.dir_x[s] = static_cast<float>(dir.x);
        t.dir_y[s] = static_cast<float>(dir.y);
        t.dir_z[s] = static_cast<float>(dir.z);
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
    t.count += count;
    t.next_id += static_cast<uint64_t>(count);
}

template <class T>
void tau_compact_array(std::vector<T>& a, const std::vector<int32_t>& src, int32_t live) {
    std::vector<T> out(live);
    for (int32_t k = 0; k < live; ++k) out[k] = a[src[k]];
    std::copy(out.begin(), out.end(), a.begin());
}

// predation_step (src/sim/predation.cu:435-490).
void predation_step(oracle_world* w, double dt) {
    oracle_world::Tau& t = w->tau;
    const int32_t n = t.count;
    if (n <= 0) return;
    const astro_world_desc& d = w->desc;
    const double dt_bio = dt * w->biology_rate;
    std::fill(w->claim.begin(), w->claim.end(), ~0ull);

    // hunt_kernel (:78-159)
    for (int32_t i = 0; i < n; ++i) {
        if (!live_flags(t.flags[i])) continue;
        V3 pos{t.x[i], t.y[i], t.z[i]};
        const bool digesting = t.digest_timer[i] > 0.0f;
        int prey_count = 0;
        int32_t prey = -1;
        for_each_neighbour(w, pos.x, pos.y, pos.z, [&](int32_t j) {
            if (!live_flags(w->flags[j])) return;
            ++prey_count;
            if (prey < 0 && tau_overlaps_cell(pos, pos_of(w, j))) prey = j;
        });
        if (!digesting && prey >= 0) {
            t.target[i] = prey;
            w->claim[prey] = std::min(w->claim[prey], t.id[i]);
        } else {
            t.target[i] = -1;
        }
        Rng rng = cell_rng(t.rng_state[i], t.id[i]);
        float ema = t.density_ema[i];
        float run = t.run_timer[i];
        V3 dir{t.dir_x[i], t.dir_y[i], t.dir_z[i]};
        if (tau_should_tumble(static_cast<float>(prey_count), ema, run, dt)) {
            dir = random_heading(rng);
            run = 0.0f;
        } else {
            run += static_cast<float>(dt);
        }
        t.density_ema[i] = ema;
        t.run_timer[i] = run;
        t.dir_x[i] = static_cast<float>(dir.x);
        t.dir_y[i] = static_cast<float>(dir.y);
        t.dir_z[i] = static_cast<float>(dir.z);
        // integrate_cell (integrator.cuh:154-166) with the crawl thrust
        V3 vel{t.vx[i], t.vy[i], t.vz[i]};
        const double t_amb = d.ambient_temp;
        const double gamma = tau_drag(t_amb);
        const Ou c = ou_coefficients(ASTRO_TAU_MASS, gamma, t_amb, dt);
        double z1, z2, z3, z4, z5, z6;
        gaussian_pair(rng, z1, z2);
        gaussian_pair(rng, z3, z4);
        gaussian_pair(rng, z5, z6);
        const bool noise = d.thermal_noise != 0;
        ou_advance_axis(pos.x, vel.x, dir.x * ASTRO_TAU_CRAWL_THRUST, gamma, c, dt, noise, z1, z2);
        ou_advance_axis(pos.y, vel.y, dir.y * ASTRO_TAU_CRAWL_THRUST, gamma, c, dt, noise, z3, z4);
        ou_advance_axis(pos.z, vel.z, dir.z * ASTRO_TAU_CRAWL_THRUST, gamma, c, dt, noise, z5, z6);
        t.rng_state[i] = rng.state;
        const double a = ASTRO_TAU_RADIUS;
        apply_boundary_axis(pos.x, vel.x, 0.5 * d.chamber_w, a, d.boundary_x);
        apply_boundary_axis(pos.y, vel.y, 0.5 * d.chamber_h, a, d.boundary_y);
        apply_boundary_axis(pos.z, vel.z, 0.5 * d.chamber_d, a, ASTRO_BOUNDARY_REFLECTING);
        t.x[i] = pos.x; t.y[i] = pos.y; t.z[i] = pos.z;
        t.vx[i] = vel.x; t.vy[i] = vel.y; t.vz[i] = vel.z;
        if (digesting) {
            float dg = t.digest_timer[i] - static_cast<float>(dt_bio);
            if (dg <= 0.0f) {
                t.biomass[i] += t.prey_biomass[i] * ASTRO_TAU_BIOMASS_YIELD;
                t.prey_biomass[i] = 0.0;
                dg = 0.0f;
            }
            t.digest_timer[i] = dg;
        }
    }

    // resolve_kernel (:163-180)
    for (int32_t i = 0; i < n; ++i) {
        const int32_t prey = t.target[i];
        if (prey < 0 || w->claim[prey] != t.id[i]) continue;
        const uint32_t cf = w->flags[prey];
        if (!live_flags(cf)) continue;
        w->flags[prey] = cf & ~ASTRO_FLAG_ALIVE;
        w->death_cause[prey] = ASTRO_DEATH_PREDATED;
        w->emit_power[prey] = 0.0f;
        w->energy[prey] = d.store_disposition == ASTRO_STORE_RETAIN ? w->energy[prey] : 0.0;   // corpse_energy
        t.digest_timer[i] = static_cast<float>(ASTRO_TAU_DIGEST_TIME);
        t.prey_biomass[i] = w->biomass[prey];
    }

    // nitrogen_kernel (:187-202)
    const Grid& N = w->f[ASTRO_FIELD_N2];
    for (int32_t i = 0; i < n; ++i) {
        const uint32_t f = t.flags[i];
        if (!live_flags(f)) continue;
        Rng rng = cell_rng(t.rng_state[i], t.id[i]);
        const double u = uniform01d(rng);
        t.rng_state[i] = rng.state;
        const double n2 = N.enabled ? static_cast<double>(grid_sample(N, static_cast<float>(t.x[i]),
                                                                       static_cast<float>(t.y[i]),
                                                                       static_cast<float>(t.z[i])))
                                    : 0.0;
        const double hazard = tau_n2_hazard(n2, static_cast<double>(t.tolerance[i]));
        if (u < tau_n2_death_prob(hazard, dt_bio)) t.flags[i] = f & ~ASTRO_FLAG_ALIVE;
    }

    // tau_mark_kernel, scan and tau_divide_kernel (:208-280, :463-481)
    std::vector<int32_t> prefix(n);
    int32_t births = 0, deads = 0;
    for (int32_t i = 0; i < n; ++i) {
        const uint32_t f = t.flags[i];
        const bool occ = (f & ASTRO_FLAG_OCCUPIED) != 0u, alive = (f & ASTRO_FLAG_ALIVE) != 0u;
        prefix[i] = births;
        if (occ && alive && t.biomass[i] >= ASTRO_TAU_DIVIDE_BIOMASS) ++births;
        if (occ && !alive) ++deads;
    }
    if (births > 0) {
        if (n + births > t.capacity) births = t.capacity - n;
        for (int32_t i = 0; births > 0 && i < n; ++i) {
            if (!live_flags(t.flags[i]) || !(t.biomass[i] >= ASTRO_TAU_DIVIDE_BIOMASS)) continue;
            const int32_t s = n + prefix[i];
            if (s >= t.capacity) continue;
            const uint64_t did = t.next_id + static_cast<uint64_t>(prefix[i]);
            const double half = 0.5 * t.biomass[i];
            t.biomass[i] = half;
            t.biomass[s] = half;
            Rng dr = cell_rng(pcg_split(t.rng_state[i], did), did);
            double mutation, unused;
            gaussian_pair(dr, mutation, unused);
            t.tolerance[s] = tau_daughter_tolerance(t.tolerance[i], mutation);
            t.rng_state[s] = dr.state;
            t.generation[s] = t.generation[i] + 1u;
            double ax, ay, az;
            division_axis(did, ax, ay, az);
            const double a = ASTRO_TAU_RADIUS;
            t.x[s] = t.x[i] + a * ax; t.x[i] -= a * ax;
            t.y[s] = t.y[i] + a * ay; t.y[i] -= a * ay;
            t.z[s] = t.z[i] + a * az; t.z[i] -= a * az;
            const double hx = 0.5 * d.chamber_w - a, hy = 0.5 * d.chamber_h - a, hz = 0.5 * d.chamber_d - a;
            auto cl = [](double v, double hh) { return v < -hh ? -hh : (v > hh ? hh : v); };
            t.x[s] = cl(t.x[s], hx); t.y[s] = cl(t.y[s], hy); t.z[s] = cl(t.z[s], hz);
            t.x[i] = cl(t.x[i], hx); t.y[i] = cl(t.y[i], hy); t.z[i] = cl(t.z[i], hz);
            t.id[s] = did;
            t.flags[s] = t.flags[i];
            t.vx[s] = t.vx[i]; t.vy[s] = t.vy[i]; t.vz[s] = t.vz[i];
            t.dir_x[s] = t.dir_x[i]; t.dir_y[s] = t.dir_y[i]; t.dir_z[s] = t.dir_z[i];
            t.prey_biomass[s] = 0.0;
            t.digest_timer[s] = 0.0f;
            t.density_ema[s] = 0.0f;
            t.run_timer[s] = 0.0f;
            t.target[s] = -1;
        }
        if (births > 0) {
            t.next_id += static_cast<uint64_t>(births);
            t.count += births;
        }
    }

    // taumoeba_store_compact (:492-527)
    if (d.tau_compaction_enabled && deads > 0) {
        const int32_t cnt = t.count;
        std::vector<int32_t> src;
        for (int32_t i = 0; i < cnt; ++i)
            if (live_flags(t.flags[i])) src.push_back(i);
        const int32_t live = static_cast<int32_t>(src.size());
        if (live < cnt) {
            tau_compact_array(t.id, src, live); tau_compact_array(t.flags, src, live);
            tau_compact_array(t.x, src, live); tau_compact_array(t.y, src, live); tau_compact_array(t.z, src, live);
            tau_compact_array(t.vx, src, live); tau_compact_array(t.vy, src, live); tau_compact_array(t.vz, src, live);
            tau_compact_array(t.dir_x, src, live); tau_compact_array(t.dir_y, src, live);
            tau_compact_array(t.dir_z, src, live); tau_compact_array(t.biomass, src, live);
            tau_compact_array(t.prey_biomass, src, live); tau_compact_array(t.digest_timer, src, live);
            tau_compact_array(t.tolerance, src, live); tau_compact_array(t.generation, src, live);
            tau_compact_array(t.rng_state, src, live); tau_compact_array(t.density_ema, src, live);
            tau_compact_array(t.run_timer, src, live); tau_compact_array(t.target, src, live);
            t.count = live;
        }
    }
}

// flash_step / flash_kernel (src/sim/flash.cu:31-67).
void flash_step(oracle_world* w, double dt) {
    for (int32_t i = 0; i < w->count; ++i) {
        const uint32_t f = w->flags[i];
        if (!(f & ASTRO_FLAG_OCCUPIED) || !(f & ASTRO_FLAG_ALIVE) || !(f & ASTRO_FLAG_AWAKE)) continue;
        const double e = w->energy[i];
        w->emit_power[i] = 0.0f;
        if (e <= 0.0) continue;
        double de = w->desc.petrova_flash_power * dt;
        if (de > e) de = e;
        w->energy[i] = e - de;
        w->flash_accum[2] += static_cast<uint64_t>(to_fixed(-(de / ASTRO_C_LIGHT), ASTRO_FLASH_IMPULSE_SCALE));
        w->flash_accum[3] += static_cast<uint64_t>(to_fixed(de, ASTRO_FLASH_ENERGY_SCALE));
    }
}

// world_step (src/sim/step.cu:109-184).
int world_step_once(oracle_world* w) {
    const double dt = ASTRO_DT_PHYSICS * w->physics_rate;
    w->cell_steps += w->count;
    hash_build(w);
    const astro_world_desc& d = w->desc;
    int st;
    if (d.thermal_enabled && (st = thermal_step(w, dt)) != ASTRO_OK) return st;
    if (d.emission_enabled && (st = emission_step(w, dt)) != ASTRO_OK) return st;
    if (w->flash_active) flash_step(w, dt);           // stage 9b, step.cu:158-162
    else if (d.taxis_enabled) taxis_step(w, dt);
    field_sample(w);
    forces(w);
    integrate(w, dt);
    if ((st = slow_fields(w, dt)) != ASTRO_OK) return st;
    predation_step(w, dt);
    if (d.lifecycle_enabled) lifecycle_step(w, dt);
    ++w->tick;
    w->sim_time += dt;
    w->hash_valid = false;
    return ASTRO_OK;
}

template <class T>
void put(std::vector<T>& dst, const T* src, int32_t n) {
    if (src) std::copy(src, src + n, dst.begin());
}
template <class T>
void get(const std::vector<T>& src, T* dst, int32_t n) {
    if (dst) std::copy(src.begin(), src.begin() + n, dst);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

// Test-only switch, per world: with it on, an FP32-layout world rounds
// positions and velocities to float where the device layout stores them (the
// integrator and division), so the oracle reproduces that layout's storage.
// Off by default: the oracle then keeps them in double in both layouts.
int32_t oracle_world_set_float_store(oracle_world* w, int32_t on) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    w->float_store = on != 0;
    return ASTRO_OK;
}

void oracle_world_desc_default(astro_world_desc* d) {
    std::memset(d, 0, sizeof(*d));
    d->abi_version = ASTRO_ABI_VERSION;
    d->capacity = 0;
    d->seed = 20260802ull;
    d->chamber_w = ASTRO_CHAMBER_W;
    d->chamber_h = ASTRO_CHAMBER_H;
    d->chamber_d = ASTRO_CHAMBER_D;
    d->precision = ASTRO_FP64_CONTRACT;
    d->hash_cell_size = 0.0;
    d->boundary_x = d->boundary_y = ASTRO_BOUNDARY_REFLECTING;
    d->gravity_axis = ASTRO_GRAVITY_Y;
    d->thermal_noise = d->contact_enabled = d->adhesion_enabled = 1;
    d->thermal_enabled = d->emission_enabled = d->taxis_enabled = d->occlusion_exact = 1;
    d->store_disposition = ASTRO_STORE_VOID;
    d->compaction_enabled = 0;
    d->lifecycle_enabled = 1;
    d->ambient_temp = ASTRO_AMBIENT_TEMP_DEFAULT;
    d->co2_init = 0.0;
    d->division_offset = 0.0;
    d->physics_rate = 1.0;
    d->biology_rate = 1.0;
    d->petrova_max_power = ASTRO_PETROVA_MAX_POWER;
    d->co2_mass_per_division = ASTRO_CO2_MASS_PER_DIVISION;
    d->petrova_flash_power = ASTRO_PETROVA_FLASH_POWER;
    d->light_dir_x = 1.0f;
    d->light_dir_y = 0.0f;
    d->light_irradiance = 0.0f;
    d->ambient_irradiance = 0.0f;
    d->light_enabled = 0;
}

const char* oracle_status_str(int32_t s) {
    static const char* names[] = {"Ok", "OutOfMemory", "DeviceError", "InvalidArgument", "NotFound",
                                  "ParseError", "ContractVersionMismatch", "CapacityExceeded",
                                  "FileIoError", "Unsupported"};
    return (s >= 0 && s <= 9) ? names[s] : "Unknown";
}

const char* oracle_last_error(const oracle_world* w) { return w ? w->err.c_str() : ""; }

int32_t oracle_world_create(const astro_world_desc* din, oracle_world** out) {
    if (!din || !out) return ASTRO_INVALID_ARGUMENT;
    if (din->abi_version != ASTRO_ABI_VERSION) return ASTRO_CONTRACT_VERSION_MISMATCH;
    astro_world_desc d = *din;
    const int32_t cap = d.capacity > 0 ? d.capacity : ASTRO_DEFAULT_CELLS;
    if (cap <= 0 || cap > ASTRO_MAX_CELLS) return ASTRO_INVALID_ARGUMENT;
    if (!(d.chamber_w > 0 && d.chamber_h > 0 && d.chamber_d > 0)) return ASTRO_INVALID_ARGUMENT;
    if (d.rng_kind != ASTRO_RNG_PCG32 && d.rng_kind != ASTRO_RNG_PHILOX) return ASTRO_INVALID_ARGUMENT;
    if (!(d.petrova_flash_power >= 0.0)) return ASTRO_INVALID_ARGUMENT;
    d.capacity = cap;
    const int32_t tau_cap = d.tau_capacity > 0 ? d.tau_capacity : ASTRO_DEFAULT_TAUMOEBA;   // step.cu:50
    if (tau_cap > ASTRO_MAX_TAUMOEBA) return ASTRO_INVALID_ARGUMENT;
    d.tau_capacity = tau_cap;
    if (d.division_offset <= 0.0) d.division_offset = ASTRO_CELL_RADIUS;
    const int defaults[4] = {ASTRO_FIELD_N_TEMP, ASTRO_FIELD_N_CO2, ASTRO_FIELD_N_N2, ASTRO_FIELD_N_IRRAD};
    for (int f = 0; f < 4; ++f) {
        astro_grid_dims& g = d.grid[f];
        if (g.nx == 0 && g.ny == 0 && g.nz == 0) { g.nx = g.ny = defaults[f]; g.nz = 1; }
        if (f == ASTRO_FIELD_N2 && g.nx < 0) continue;   // disabled
        if (g.nx <= 1 || g.ny <= 1 || g.nz < 1) return ASTRO_INVALID_ARGUMENT;
    }
    auto* w = new oracle_world();
    resize_tau(w->tau, tau_cap);
    w->claim.assign(cap, ~0ull);
    w->desc = d;
    w->capacity = cap;
    w->physics_rate = d.physics_rate > 0 ? d.physics_rate : 1.0;
    w->biology_rate = d.biology_rate > 0 ? d.biology_rate : 1.0;
    resize_cells(w, cap);
    const double W = d.chamber_w, H = d.chamber_h, Dd = d.chamber_d;
    grid_init(w->f[ASTRO_FIELD_TEMPERATURE], d.grid[0], W, H, Dd, ASTRO_WATER_THERMAL_DIFFUSIVITY,
              d.ambient_temp, ASTRO_DEPOSIT_SCALE_TEMPERATURE, ASTRO_BC_ROBIN);
    grid_init(w->f[ASTRO_FIELD_CO2], d.grid[1], W, H, Dd, ASTRO_CO2_DIFFUSIVITY_WATER, d.co2_init,
              ASTRO_DEPOSIT_SCALE_CO2, ASTRO_BC_NEUMANN);
    grid_init(w->f[ASTRO_FIELD_N2], d.grid[2], W, H, Dd, ASTRO_N2_DIFFUSIVITY, 0.0,
              ASTRO_DEPOSIT_SCALE_N2, ASTRO_BC_NEUMANN);
    grid_init(w->f[ASTRO_FIELD_IRRADIANCE], d.grid[3], W, H, Dd, ASTRO_IRRADIANCE_NOMINAL_DIFFUSIVITY,
              0.0, ASTRO_DEPOSIT_SCALE_IRRADIANCE, ASTRO_BC_NEUMANN);
    w->co2_demand.assign(w->f[ASTRO_FIELD_CO2].count(), 0ull);
    // hash_create (src/sim/hash.cu:46-80) and hash_view (:90-104)
    w->hcell = d.hash_cell_size > 0.0 ? d.hash_cell_size : ASTRO_HASH_CELL_FACTOR * ASTRO_CELL_DIAMETER;
    w->hnx = std::max(1, static_cast<int32_t>(std::ceil(W / w->hcell)));
    w->hny = std::max(1, static_cast<int32_t>(std::ceil(H / w->hcell)));
    w->hnz = std::max(1, static_cast<int32_t>(std::ceil(Dd / w->hcell)));
    w->nb = w->hnx * w->hny * w->hnz;
    w->h_inv = static_cast<float>(1.0 / w->hcell);
    w->h_ox = static_cast<float>(-0.5 * w->hnx * w->hcell);
    w->h_oy = static_cast<float>(-0.5 * w->hny * w->hcell);
    w->h_oz = static_cast<float>(-0.5 * w->hnz * w->hcell);
    w->bstart.assign(w->nb, 0);
    w->bend.assign(w->nb, 0);
    *out = w;
    return ASTRO_OK;
}

void oracle_world_destroy(oracle_world* w) { delete w; }

int32_t oracle_world_desc_get(const oracle_world* w, astro_world_desc* out) {
    if (!w || !out) return ASTRO_INVALID_ARGUMENT;
    *out = w->desc;
    return ASTRO_OK;
}

// spawn_kernel (src/sim/cell_store.cu:42-130) + cell_store_spawn (:262-277).
int32_t oracle_cells_spawn(oracle_world* w, const astro_spawn_params* p, uint64_t seed) {
    if (!w || !p) return ASTRO_INVALID_ARGUMENT;
    if (p->count <= 0) return ASTRO_OK;
    if (w->count + p->count > w->capacity) return fail(w, ASTRO_CAPACITY_EXCEEDED, "spawn exceeds cell store capacity");
    const astro_world_desc& d = w->desc;
    for (int32_t i = 0; i < p->count; ++i) {
        const int32_t slot = w->count + i;
        const uint64_t id = w->next_id + static_cast<uint64_t>(i);
        const bool philox = d.rng_kind == ASTRO_RNG_PHILOX;
        Rng r = philox ? philox_stream(seed, ASTRO_RNG_STAGE_SPAWN, 0u, id) : cell_rng(cell_rng_init(seed, id), id);
        double x = 0.0, y = 0.0;
        switch (p->placement) {
            case ASTRO_PLACE_GAUSSIAN: {
                double gx, gy;
                gaussian_pair(r, gx, gy);
                x = p->place_x + p->place_radius * gx;
                y = p->place_y + p->place_radius * gy;
                break;
            }
            case ASTRO_PLACE_DISC: {
                const double rad = p->place_radius * std::sqrt(uniform01d(r));
                const double ang = uniform_range(r, 0.0, ASTRO_TWO_PI);
                x = p->place_x + rad * std::cos(ang);
                y = p->place_y + rad * std::sin(ang);
                break;
            }
            case ASTRO_PLACE_GRID: {
                const int side = static_cast<int>(std::ceil(std::sqrt(static_cast<double>(p->count))));
                const int gx = i % side, gy = i / side;
                x = (static_cast<double>(gx) / side - 0.5) * d.chamber_w * 0.98;
                y = (static_cast<double>(gy) / side - 0.5) * d.chamber_h * 0.98;
                break;
            }
            default:
                x = uniform_range(r, -0.5 * d.chamber_w, 0.5 * d.chamber_w);
                y = uniform_range(r, -0.5 * d.chamber_h, 0.5 * d.chamber_h);
                break;
        }
        const double a = ASTRO_CELL_RADIUS;
        auto cl = [](double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); };
        x = cl(x, -0.5 * d.chamber_w + a, 0.5 * d.chamber_w - a);
        y = cl(y, -0.5 * d.chamber_h + a, 0.5 * d.chamber_h - a);
        const double z = uniform_range(r, -0.5 * d.chamber_d + a, 0.5 * d.chamber_d - a);
        double charge;
        switch (p->charge_dist) {
            case ASTRO_DIST_UNIFORM: charge = uniform_range(r, p->charge_a, p->charge_b); break;
            case ASTRO_DIST_NORMAL: {
                double z0, z1;
                gaussian_pair(r, z0, z1);
                charge = p->charge_a + p->charge_b * z0;
                break;
            }
            case ASTRO_DIST_LOG_UNIFORM:
                charge = std::exp(uniform_range(r, std::log(p->charge_a), std::log(p->charge_b)));
                break;
            default: charge = p->charge_a; break;
        }
        charge = cl(charge, 0.0, 1.0);
        w->id[slot] = id;
        w->flags[slot] = ASTRO_FLAG_OCCUPIED | ASTRO_FLAG_ALIVE | (p->awake ? ASTRO_FLAG_AWAKE : 0u);
        w->death_cause[slot] = ASTRO_DEATH_NONE;
        w->x[slot] = x; w->y[slot] = y; w->z[slot] = z;
        w->vx[slot] = w->vy[slot] = w->vz[slot] = 0.0;
        w->energy[slot] = charge * ASTRO_CELL_ENERGY_MAX;
        w->temp_cell[slot] = static_cast<float>(p->awake ? ASTRO_CELL_TEMP_SETPOINT : ASTRO_AMBIENT_TEMP_DEFAULT);
        w->emit_power[slot] = 0.0f;
        double ga, gb, gc, gd;                                   // gaussian3, rng.cuh:111-117
        gaussian_pair(r, ga, gb);
        gaussian_pair(r, gc, gd);
        const V3 dir = normalize(V3{ga, gb, gc});
        w->dir_x[slot] = static_cast<float>(dir.x);
        w->dir_y[slot] = static_cast<float>(dir.y);
        w->dir_z[slot] = static_cast<float>(dir.z);
        w->biomass[slot] = ASTRO_CELL_MASS_DRY;
        w->co2_held[slot] = 0.0;
        w->age_s[slot] = 0.0f;
        w->rng_state[slot] = philox ? 0ull : r.state;
        w->taxis_memory[slot] = 0.0f;
        w->run_timer[slot] = 0.0f;
        w->t_local[slot] = w->temp_cell[slot];
        w->irradiance[slot] = 0.0f;
        w->co2_local[slot] = 0.0f;
        w->n2_local[slot] = 0.0f;
    }
    w->count += p->count;
    w->next_id += static_cast<uint64_t>(p->count);
    w->hash_valid = false;
    return ASTRO_OK;
}

int32_t oracle_cells_upload(oracle_world* w, const astro_cell_soa* s, int32_t n) {
    if (!w || !s || n < 0) return ASTRO_INVALID_ARGUMENT;
    if (n > w->capacity) return fail(w, ASTRO_CAPACITY_EXCEEDED, "upload exceeds capacity");
    put(w->id, s->id, n); put(w->flags, s->flags, n); put(w->death_cause, s->death_cause, n);
    put(w->x, s->x, n); put(w->y, s->y, n); put(w->z, s->z, n);
    put(w->vx, s->vx, n); put(w->vy, s->vy, n); put(w->vz, s->vz, n);
    put(w->energy, s->energy, n); put(w->temp_cell, s->temp_cell, n);
    put(w->emit_power, s->emit_power, n);
    put(w->dir_x, s->dir_x, n); put(w->dir_y, s->dir_y, n); put(w->dir_z, s->dir_z, n);
    put(w->biomass, s->biomass, n); put(w->co2_held, s->co2_held, n); put(w->age_s, s->age_s, n);
    put(w->rng_state, s->rng_state, n); put(w->taxis_memory, s->taxis_memory, n);
    put(w->run_timer, s->run_timer, n); put(w->t_local, s->t_local, n);
    put(w->irradiance, s->irradiance, n); put(w->co2_local, s->co2_local, n);
    put(w->n2_local, s->n2_local, n);
    w->count = n;
    w->hash_valid = false;
    return ASTRO_OK;
}

int32_t oracle_cells_download(oracle_world* w, astro_cell_soa* s, int32_t n, int32_t* n_out) {
    if (!w || !s) return ASTRO_INVALID_ARGUMENT;
    const int32_t m = std::min(n, w->count);
    get(w->id, s->id, m); get(w->flags, s->flags, m); get(w->death_cause, s->death_cause, m);
    get(w->x, s->x, m); get(w->y, s->y, m); get(w->z, s->z, m);
    get(w->vx, s->vx, m); get(w->vy, s->vy, m); get(w->vz, s->vz, m);
    get(w->energy, s->energy, m); get(w->temp_cell, s->temp_cell, m);
    get(w->emit_power, s->emit_power, m);
    get(w->dir_x, s->dir_x, m); get(w->dir_y, s->dir_y, m); get(w->dir_z, s->dir_z, m);
    get(w->biomass, s->biomass, m); get(w->co2_held, s->co2_held, m); get(w->age_s, s->age_s, m);
    get(w->rng_state, s->rng_state, m); get(w->taxis_memory, s->taxis_memory, m);
    get(w->run_timer, s->run_timer, m); get(w->t_local, s->t_local, m);
    get(w->irradiance, s->irradiance, m); get(w->co2_local, s->co2_local, m);
    get(w->n2_local, s->n2_local, m);
    if (n_out) *n_out = m;
    return ASTRO_OK;
}

int32_t oracle_world_get_counters(oracle_world* w, astro_counters* o) {
    if (!w || !o) return ASTRO_INVALID_ARGUMENT;
    std::memset(o, 0, sizeof(*o));
    o->count = w->count; o->capacity = w->capacity; o->next_id = w->next_id; o->tick = w->tick;
    o->sim_time_s = w->sim_time; o->divisions_total = w->divisions_total;
    o->deaths_total = w->deaths_total; o->dead_slots_prev = w->dead_slots_prev;
    o->cell_steps = w->cell_steps;
    return ASTRO_OK;
}

int32_t oracle_world_set_counters(oracle_world* w, const astro_counters* c) {
    if (!w || !c) return ASTRO_INVALID_ARGUMENT;
    if (c->count < 0 || c->count > w->capacity) return ASTRO_CAPACITY_EXCEEDED;
    w->count = c->count; w->next_id = c->next_id; w->tick = c->tick; w->sim_time = c->sim_time_s;
    w->divisions_total = c->divisions_total; w->deaths_total = c->deaths_total;
    w->dead_slots_prev = c->dead_slots_prev;
    w->cell_steps = c->cell_steps;
    w->hash_valid = false;
    return ASTRO_OK;
}

int32_t oracle_field_dims(const oracle_world* w, int32_t f, astro_grid_dims* o, double* dx,
                          double* dy, double* dz) {
    if (!w || f < 0 || f >= ASTRO_FIELD_COUNT) return ASTRO_INVALID_ARGUMENT;
    const Grid& g = w->f[f];
    if (o) { o->nx = g.nx; o->ny = g.ny; o->nz = g.nz; }
    if (dx) *dx = g.dx;
    if (dy) *dy = g.dy;
    if (dz) *dz = g.dz;
    return ASTRO_OK;
}

int32_t oracle_field_upload(oracle_world* w, int32_t f, const float* v, size_t count) {
    if (!w || f < 0 || f >= ASTRO_FIELD_COUNT || !v) return ASTRO_INVALID_ARGUMENT;
    Grid& g = w->f[f];
    if (!g.enabled || count != g.count()) return fail(w, ASTRO_INVALID_ARGUMENT, "field size mismatch");
    std::copy(v, v + count, g.value.begin());
    std::copy(v, v + count, g.scratch.begin());
    return ASTRO_OK;
}

int32_t oracle_field_download(oracle_world* w, int32_t f, float* v, size_t count) {
    if (!w || f < 0 || f >= ASTRO_FIELD_COUNT || !v) return ASTRO_INVALID_ARGUMENT;
    const Grid& g = w->f[f];
    if (!g.enabled || count != g.count()) return fail(w, ASTRO_INVALID_ARGUMENT, "field size mismatch");
    std::copy(g.value.begin(), g.value.end(), v);
    return ASTRO_OK;
}

int32_t oracle_field_deposit_download(oracle_world* w, int32_t f, int64_t* v, size_t count) {
    if (!w || f < 0 || f >= ASTRO_FIELD_COUNT || !v) return ASTRO_INVALID_ARGUMENT;
    const Grid& g = w->f[f];
    if (!g.enabled || count != g.count()) return fail(w, ASTRO_INVALID_ARGUMENT, "field size mismatch");
    for (size_t i = 0; i < count; ++i) v[i] = static_cast<int64_t>(g.deposit[i]);
    return ASTRO_OK;
}

int32_t oracle_world_step(oracle_world* w, int32_t n_ticks) {
    if (!w || n_ticks < 0) return ASTRO_INVALID_ARGUMENT;
    for (int32_t t = 0; t < n_ticks; ++t) {
        const int st = world_step_once(w);
        if (st != ASTRO_OK) return st;
    }
    return ASTRO_OK;
}

int32_t oracle_world_sync(oracle_world* w) { return w ? ASTRO_OK : ASTRO_INVALID_ARGUMENT; }

// world_set_clock with a preset (src/sim/step.cu:190-215).
int32_t oracle_world_set_clock_preset(oracle_world* w, int32_t preset, double physics, double biology) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    if (preset == ASTRO_CLOCK_MOTION) {
        physics = ASTRO_CLOCK_MOTION_PHYSICS; biology = ASTRO_CLOCK_MOTION_BIOLOGY;
    } else if (preset == ASTRO_CLOCK_METABOLIC) {
        physics = ASTRO_CLOCK_METABOLIC_PHYSICS; biology = ASTRO_CLOCK_METABOLIC_BIOLOGY;
    } else if (preset == ASTRO_CLOCK_GENERATIONAL) {
        physics = ASTRO_CLOCK_GENERATIONAL_PHYSICS; biology = ASTRO_CLOCK_GENERATIONAL_BIOLOGY;
    } else if (preset != ASTRO_CLOCK_CUSTOM) {
        physics = ASTRO_CLOCK_REALTIME_PHYSICS; biology = ASTRO_CLOCK_REALTIME_BIOLOGY;
    }
    return oracle_world_set_clock(w, physics, biology);
}

int32_t oracle_world_set_light(oracle_world* w, float dir_x, float dir_y, float irradiance, float ambient,
                               int32_t enabled) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    w->desc.light_dir_x = dir_x;
    w->desc.light_dir_y = dir_y;
    w->desc.light_irradiance = irradiance;
    w->desc.ambient_irradiance = ambient;
    w->desc.light_enabled = enabled ? 1 : 0;
    return ASTRO_OK;
}

// world_apply_brush (src/sim/step.cu:80-89) and brush_kernel
// (src/fields/diffuse.cu:62-89), per voxel column in float.
int32_t oracle_world_apply_brush(oracle_world* w, int32_t kind, double x, double y, double radius,
                                 double strength) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    int f = -1;
    double amount = strength;
    if (kind == ASTRO_BRUSH_HEAT) f = ASTRO_FIELD_TEMPERATURE;
    else if (kind == ASTRO_BRUSH_CHILL) { f = ASTRO_FIELD_TEMPERATURE; amount = -strength; }
    else if (kind == ASTRO_BRUSH_INJECT_CO2) f = ASTRO_FIELD_CO2;
    else if (kind == ASTRO_BRUSH_INJECT_N2) f = ASTRO_FIELD_N2;
    else return fail(w, ASTRO_INVALID_ARGUMENT, "unknown brush");
    Grid& g = w->f[f];
    if (!g.enabled) return fail(w, ASTRO_INVALID_ARGUMENT, "brush on a disabled field");
    if (radius <= 0.0) return ASTRO_OK;
    const float cx = static_cast<float>(x), cy = static_cast<float>(y);
    const float rad = static_cast<float>(radius), am = static_cast<float>(amount);
    const float halfx = 0.5f * static_cast<float>(g.nx) * g.fdx;
    const float halfy = 0.5f * static_cast<float>(g.ny) * g.fdy;
    for (int k = 0; k < g.nz; ++k)
        for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i < g.nx; ++i) {
                const float px = (static_cast<float>(i) + 0.5f) * g.fdx - halfx;
                const float py = (static_cast<float>(j) + 0.5f) * g.fdy - halfy;
                const float r = std::sqrt((px - cx) * (px - cx) + (py - cy) * (py - cy));
                if (r > rad) continue;
                const float t = r / rad;
                const float wgt = (1.0f - t * t) * (1.0f - t * t);
                float& v = g.value[g.at(i, j, k)];
                v = v + am * wgt;
            }
    return ASTRO_OK;
}

// grid_fill (src/fields/diffuse.cu:128-134).
int32_t oracle_field_fill(oracle_world* w, int32_t f, float value) {
    if (!w || f < 0 || f >= ASTRO_FIELD_COUNT) return ASTRO_INVALID_ARGUMENT;
    Grid& g = w->f[f];
    if (!g.enabled) return fail(w, ASTRO_INVALID_ARGUMENT, "fill of a disabled field");
    std::fill(g.value.begin(), g.value.end(), value);
    std::fill(g.scratch.begin(), g.scratch.end(), value);
    return ASTRO_OK;
}

int32_t oracle_world_set_flash(oracle_world* w, int32_t armed) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    w->flash_active = armed != 0;
    return ASTRO_OK;
}

// world_flash_audit (src/sim/flash.cu:69-79).
int32_t oracle_world_flash_audit(oracle_world* w, double* impulse_ns, double* discharged_j) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    const double ix = from_fixed(w->flash_accum[0], ASTRO_FLASH_IMPULSE_SCALE);
    const double iy = from_fixed(w->flash_accum[1], ASTRO_FLASH_IMPULSE_SCALE);
    const double iz = from_fixed(w->flash_accum[2], ASTRO_FLASH_IMPULSE_SCALE);
    if (impulse_ns) *impulse_ns = std::sqrt(ix * ix + iy * iy + iz * iz);
    if (discharged_j) *discharged_j = from_fixed(w->flash_accum[3], ASTRO_FLASH_ENERGY_SCALE);
    return ASTRO_OK;
}

int32_t oracle_world_time(const oracle_world* w, uint64_t* tick, double* sim_time_s, double* physics_rate,
                          double* biology_rate) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    if (tick) *tick = w->tick;
    if (sim_time_s) *sim_time_s = w->sim_time;
    if (physics_rate) *physics_rate = w->physics_rate;
    if (biology_rate) *biology_rate = w->biology_rate;
    return ASTRO_OK;
}

// taumoeba_spawn (src/sim/predation.cu:421-433).
int32_t oracle_taumoeba_spawn(oracle_world* w, int32_t count, uint64_t seed) {
    if (!w || count < 0) return ASTRO_INVALID_ARGUMENT;
    if (count == 0) return ASTRO_OK;
    if (w->tau.count + count > w->tau.capacity)
        return fail(w, ASTRO_CAPACITY_EXCEEDED, "taumoeba spawn exceeds capacity");
    tau_spawn(w, count, seed);
    return ASTRO_OK;
}

int32_t oracle_taumoeba_upload(oracle_world* w, const astro_tau_soa* s, int32_t n) {
    if (!w || !s || n < 0) return ASTRO_INVALID_ARGUMENT;
    if (n > w->tau.capacity) return fail(w, ASTRO_CAPACITY_EXCEEDED, "taumoeba upload exceeds capacity");
    oracle_world::Tau& t = w->tau;
    put(t.id, s->id, n); put(t.flags, s->flags, n); put(t.x, s->x, n); put(t.y, s->y, n); put(t.z, s->z, n);
    put(t.vx, s->vx, n); put(t.vy, s->vy, n); put(t.vz, s->vz, n); put(t.dir_x, s->dir_x, n);
    put(t.dir_y, s->dir_y, n); put(t.dir_z, s->dir_z, n); put(t.biomass, s->biomass, n);
    put(t.prey_biomass, s->prey_biomass, n); put(t.digest_timer, s->digest_timer, n);
    put(t.tolerance, s->tolerance, n); put(t.generation, s->generation, n); put(t.rng_state, s->rng_state, n);
    put(t.density_ema, s->density_ema, n); put(t.run_timer, s->run_timer, n); put(t.target, s->target, n);
    t.count = n;
    uint64_t mx = 0;   // src/sim/snapshot.cu:239-252
    for (int32_t i = 0; i < n; ++i) mx = std::max(mx, s->id[i]);
    t.next_id = mx + 1;
    return ASTRO_OK;
}

int32_t oracle_taumoeba_download(oracle_world* w, astro_tau_soa* s, int32_t n, int32_t* count_out) {
    if (!w || !s || n < 0) return ASTRO_INVALID_ARGUMENT;
    const oracle_world::Tau& t = w->tau;
    if (count_out) *count_out = t.count;
    const int32_t m = std::min(n, t.count);
    get(t.id, s->id, m); get(t.flags, s->flags, m); get(t.x, s->x, m); get(t.y, s->y, m); get(t.z, s->z, m);
    get(t.vx, s->vx, m); get(t.vy, s->vy, m); get(t.vz, s->vz, m); get(t.dir_x, s->dir_x, m);
    get(t.dir_y, s->dir_y, m); get(t.dir_z, s->dir_z, m); get(t.biomass, s->biomass, m);
    get(t.prey_biomass, s->prey_biomass, m); get(t.digest_timer, s->digest_timer, m);
    get(t.tolerance, s->tolerance, m); get(t.generation, s->generation, m); get(t.rng_state, s->rng_state, m);
    get(t.density_ema, s->density_ema, m); get(t.run_timer, s->run_timer, m); get(t.target, s->target, m);
    return ASTRO_OK;
}

int32_t oracle_taumoeba_counters(oracle_world* w, int32_t* count, int32_t* capacity, uint64_t* next_id) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    if (count) *count = w->tau.count;
    if (capacity) *capacity = w->tau.capacity;
    if (next_id) *next_id = w->tau.next_id;
    return ASTRO_OK;
}

int32_t oracle_stage_predation(oracle_world* w, double dt) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    if (!w->hash_valid) hash_build(w);
    predation_step(w, dt);
    return ASTRO_OK;
}

int32_t oracle_world_set_clock(oracle_world* w, double physics, double biology) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    w->physics_rate = physics < ASTRO_CLOCK_PHYSICS_MIN ? ASTRO_CLOCK_PHYSICS_MIN
                    : (physics > ASTRO_CLOCK_PHYSICS_MAX ? ASTRO_CLOCK_PHYSICS_MAX : physics);
    w->biology_rate = biology < ASTRO_CLOCK_BIOLOGY_MIN ? ASTRO_CLOCK_BIOLOGY_MIN
                    : (biology > ASTRO_CLOCK_BIOLOGY_MAX ? ASTRO_CLOCK_BIOLOGY_MAX : biology);
    w->desc.physics_rate = w->physics_rate;
    w->desc.biology_rate = w->biology_rate;
    return ASTRO_OK;
}

// stats_step + world_stats (src/sim/stats.cu:27-153).
int32_t oracle_world_stats(oracle_world* w, astro_stats* s) {
    if (!w || !s) return ASTRO_INVALID_ARGUMENT;
    // stats_taumoeba_kernel (src/sim/stats.cu:65-74)
    uint32_t n_tau = 0;
    uint64_t tau_tol = 0;
    for (int32_t i = 0; i < w->tau.count; ++i) {
        if (!live_flags(w->tau.flags[i])) continue;
        ++n_tau;
        tau_tol += static_cast<uint64_t>(to_fixed(static_cast<double>(w->tau.tolerance[i]), ASTRO_DEPOSIT_SCALE_STATS));
    }
    uint64_t energy = 0, temp_cell = 0, temp_medium = 0, co2 = 0, n2 = 0, max_temp = 0;
    uint32_t n_alive = 0, n_dead = 0, n_awake = 0, n_dormant = 0;
    for (int32_t i = 0; i < w->count; ++i) {
        const uint32_t f = w->flags[i];
        if (!(f & ASTRO_FLAG_OCCUPIED)) continue;
        const bool alive = (f & ASTRO_FLAG_ALIVE) != 0u, awake = (f & ASTRO_FLAG_AWAKE) != 0u;
        if (alive) ++n_alive; else ++n_dead;
        if (awake) ++n_awake; else ++n_dormant;
        energy += static_cast<uint64_t>(to_fixed(w->energy[i], ASTRO_DEPOSIT_SCALE_STATS));
        if (alive) temp_cell += static_cast<uint64_t>(to_fixed(static_cast<double>(w->temp_cell[i]), ASTRO_DEPOSIT_SCALE_STATS));
    }
    const Grid& T = w->f[ASTRO_FIELD_TEMPERATURE];
    for (size_t i = 0; i < T.count(); ++i) {
        const double t = static_cast<double>(T.value[i]);
        temp_medium += static_cast<uint64_t>(to_fixed(t, ASTRO_DEPOSIT_SCALE_STATS));
        const uint64_t ft = static_cast<uint64_t>(to_fixed(t, ASTRO_DEPOSIT_SCALE_STATS));
        if (ft > max_temp) max_temp = ft;
    }
    const Grid& C = w->f[ASTRO_FIELD_CO2];
    for (size_t i = 0; i < C.count(); ++i)
        co2 += static_cast<uint64_t>(to_fixed(static_cast<double>(C.value[i]), ASTRO_STATS_SCALE_FIELD));
    const Grid& N = w->f[ASTRO_FIELD_N2];
    for (size_t i = 0; N.enabled && i < N.count(); ++i)
        n2 += static_cast<uint64_t>(to_fixed(static_cast<double>(N.value[i]), ASTRO_STATS_SCALE_FIELD));

    std::memset(s, 0, sizeof(*s));
    s->tick = w->tick;
    s->sim_time_s = w->sim_time;
    s->physics_rate = w->physics_rate;
    s->biology_rate = w->biology_rate;
    s->n_live = static_cast<int32_t>(n_alive);
    s->n_dead = static_cast<int32_t>(n_dead);
    s->n_awake = static_cast<int32_t>(n_awake);
    s->n_dormant = static_cast<int32_t>(n_dormant);
    s->n_taumoeba = static_cast<int32_t>(n_tau);
    s->mean_tau_tolerance = n_tau > 0u ? from_fixed(tau_tol, ASTRO_DEPOSIT_SCALE_STATS) / n_tau : 0.0;   // stats.cu:118-120
    s->total_energy_j = from_fixed(energy, ASTRO_DEPOSIT_SCALE_STATS);
    s->mean_charge = (n_alive + n_dead) > 0u ? s->total_energy_j / ((n_alive + n_dead) * ASTRO_CELL_ENERGY_MAX) : 0.0;
    s->mean_temp_cell_k = n_alive > 0u ? from_fixed(temp_cell, ASTRO_DEPOSIT_SCALE_STATS) / n_alive : 0.0;
    const double cells = static_cast<double>(T.count());
    s->mean_temp_medium_k = cells > 0.0 ? from_fixed(temp_medium, ASTRO_DEPOSIT_SCALE_STATS) / cells : 0.0;
    s->max_temp_medium_k = from_fixed(max_temp, ASTRO_DEPOSIT_SCALE_STATS);
    const double vc = C.dx * C.dy * C.dz;
    const double vn = N.enabled ? N.dx * N.dy * N.dz : 0.0;
    s->co2_total_kg = from_fixed(co2, ASTRO_STATS_SCALE_FIELD) * vc;
    s->n2_total_kg = from_fixed(n2, ASTRO_STATS_SCALE_FIELD) * vn;
    s->divisions_this_window = static_cast<int32_t>(w->divisions_total - w->divisions_reported);
    s->deaths_this_window = static_cast<int32_t>(w->deaths_total - w->deaths_reported);
    w->divisions_reported = w->divisions_total;
    w->deaths_reported = w->deaths_total;
    s->boil_events = s->max_temp_medium_k >= 373.14999999999998 ? 1 : 0;
    s->non_canon_run = (w->desc.petrova_max_power != ASTRO_PETROVA_MAX_POWER ||
                        w->desc.petrova_flash_power != ASTRO_PETROVA_FLASH_POWER ||
                        w->desc.co2_mass_per_division != ASTRO_CO2_MASS_PER_DIVISION) ? 1 : 0;
    return ASTRO_OK;
}

// snapshot_state_hash (src/sim/snapshot.cu:100-129): overrides, cell SoA in
// contract order, then T, CO2, N2 values.
int32_t oracle_world_state_hash(oracle_world* w, uint64_t* out) {
    if (!w || !out) return ASTRO_INVALID_ARGUMENT;
    uint64_t h = 1469598103934665603ull;
    struct Ov { char key[48]; double value; };
    auto ov = [&](const char* key, double v, double canon) {
        if (v == canon) return;
        Ov o;
        std::memset(&o, 0, sizeof(o));
        std::strncpy(o.key, key, sizeof(o.key) - 1);
        o.value = v;
        h = fnv1a64(&o, sizeof(o), h);
    };
    // src/sim/snapshot.cu:79-83, in that order
    ov("PETROVA_MAX_POWER", w->desc.petrova_max_power, ASTRO_PETROVA_MAX_POWER);
    ov("PETROVA_FLASH_POWER", w->desc.petrova_flash_power, ASTRO_PETROVA_FLASH_POWER);
    ov("CO2_MASS_PER_DIVISION", w->desc.co2_mass_per_division, ASTRO_CO2_MASS_PER_DIVISION);
    const size_t n = static_cast<size_t>(w->count);
    auto span = [&](const auto& v) {
        if (n > 0) h = fnv1a64(v.data(), sizeof(v[0]) * n, h);
    };
    span(w->id); span(w->flags); span(w->death_cause);
    span(w->x); span(w->y); span(w->z); span(w->vx); span(w->vy); span(w->vz);
    span(w->energy); span(w->temp_cell); span(w->emit_power);
    span(w->dir_x); span(w->dir_y); span(w->dir_z);
    span(w->biomass); span(w->co2_held); span(w->age_s);
    span(w->rng_state); span(w->taxis_memory); span(w->run_timer);
    span(w->t_local); span(w->irradiance); span(w->co2_local); span(w->n2_local);
    // collect_tau_spans (src/sim/snapshot.cu:59-67)
    const oracle_world::Tau& t = w->tau;
    const size_t nt = static_cast<size_t>(t.count);
    auto tspan = [&](const auto& v) {
        if (nt > 0) h = fnv1a64(v.data(), sizeof(v[0]) * nt, h);
    };
    tspan(t.id); tspan(t.flags); tspan(t.x); tspan(t.y); tspan(t.z); tspan(t.vx); tspan(t.vy); tspan(t.vz);
    tspan(t.dir_x); tspan(t.dir_y); tspan(t.dir_z); tspan(t.biomass); tspan(t.prey_biomass);
    tspan(t.digest_timer); tspan(t.tolerance); tspan(t.generation); tspan(t.rng_state);
    tspan(t.density_ema); tspan(t.run_timer); tspan(t.target);
    for (int f : {ASTRO_FIELD_TEMPERATURE, ASTRO_FIELD_CO2, ASTRO_FIELD_N2}) {
        const Grid& g = w->f[f];
        if (g.enabled && g.count() > 0) h = fnv1a64(g.value.data(), sizeof(float) * g.count(), h);
    }
    *out = h;
    return ASTRO_OK;
}

int32_t oracle_stage_hash_build(oracle_world* w) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    hash_build(w);
    return ASTRO_OK;
}

int32_t oracle_hash_dims(const oracle_world* w, int32_t* nx, int32_t* ny, int32_t* nz, double* cell) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    if (nx) *nx = w->hnx;
    if (ny) *ny = w->hny;
    if (nz) *nz = w->hnz;
    if (cell) *cell = w->hcell;
    return ASTRO_OK;
}

int32_t oracle_hash_download(oracle_world* w, uint32_t* vals, int32_t n, int32_t* bs, int32_t* be, int32_t nb) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    if (vals) std::copy(w->vals_sorted.begin(), w->vals_sorted.begin() + std::min(n, w->count), vals);
    const int32_t m = std::min(nb, w->nb);
    if (bs) std::copy(w->bstart.begin(), w->bstart.begin() + m, bs);
    if (be) std::copy(w->bend.begin(), w->bend.begin() + m, be);
    return ASTRO_OK;
}

int32_t oracle_stage_thermal(oracle_world* w, double dt) { return w ? thermal_step(w, dt) : ASTRO_INVALID_ARGUMENT; }
int32_t oracle_stage_emission(oracle_world* w, double dt) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    if (!w->hash_valid) hash_build(w);
    return emission_step(w, dt);
}
int32_t oracle_stage_taxis(oracle_world* w, double dt) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    taxis_step(w, dt);
    return ASTRO_OK;
}
int32_t oracle_stage_forces(oracle_world* w) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    if (!w->hash_valid) hash_build(w);
    field_sample(w);
    forces(w);
    return ASTRO_OK;
}
int32_t oracle_forces_download(oracle_world* w, double* fx, double* fy, double* fz, int32_t n) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    const int32_t m = std::min(n, w->count);
    get(w->fx, fx, m); get(w->fy, fy, m); get(w->fz, fz, m);
    return ASTRO_OK;
}
int32_t oracle_stage_integrate(oracle_world* w, double dt) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    integrate(w, dt);
    w->hash_valid = false;
    return ASTRO_OK;
}
int32_t oracle_stage_motion(oracle_world* w, double dt) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    if (!w->hash_valid) hash_build(w);
    field_sample(w);
    forces(w);
    integrate(w, dt);
    w->hash_valid = false;
    return ASTRO_OK;
}
int32_t oracle_stage_slow_fields(oracle_world* w, double dt) { return w ? slow_fields(w, dt) : ASTRO_INVALID_ARGUMENT; }
int32_t oracle_stage_diffuse(oracle_world* w, int32_t f, double dt) {
    if (!w || f < 0 || f >= ASTRO_FIELD_COUNT) return ASTRO_INVALID_ARGUMENT;
    return grid_diffuse(w->f[f], dt) ? ASTRO_OK : fail(w, ASTRO_INVALID_ARGUMENT, "diffusion coefficient unstable");
}
int32_t oracle_stage_lifecycle(oracle_world* w, double dt) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    lifecycle_step(w, dt);
    w->hash_valid = false;
    return ASTRO_OK;
}
int32_t oracle_stage_compact(oracle_world* w) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    compact(w);
    w->hash_valid = false;
    return ASTRO_OK;
}

// ---- building blocks for known-answer tests ----
uint64_t oracle_splitmix64(uint64_t x) { return splitmix64(x); }
void oracle_pcg32_seed(uint64_t is, uint64_t iq, uint64_t* st, uint64_t* inc) {
    const Rng r = pcg_seed(is, iq);
    *st = r.state;
    *inc = r.inc;
}
// Threads the parallel loops use: OpenMP's team size in the all-cores build,
// 1 in the single-threaded one.
int32_t oracle_threads(void) {
#ifdef _OPENMP
    return static_cast<int32_t>(omp_get_max_threads());
#else
    return 1;
#endif
}

int32_t oracle_philox_blocks(const uint32_t* words, int32_t n, uint32_t* out) {
    if (n < 0 || (n > 0 && (!words || !out))) return ASTRO_INVALID_ARGUMENT;
    for (int32_t i = 0; i < n; ++i) {
        const uint32_t* w = words + 6 * static_cast<size_t>(i);
        uint32_t c[4] = {w[0], w[1], w[2], w[3]};
        philox4x32_10(c, w[4], w[5]);
        for (int k = 0; k < 4; ++k) out[4 * static_cast<size_t>(i) + k] = c[k];
    }
    return ASTRO_OK;
}

int32_t oracle_philox_stream(uint64_t seed, uint64_t id, uint64_t tick, int32_t stage, int32_t n, uint32_t* out) {
    if (n < 0 || stage < 0 || stage > 0xFFFF || (n > 0 && !out)) return ASTRO_INVALID_ARGUMENT;
    Rng r = philox_stream(seed, static_cast<uint32_t>(stage), tick, id);
    for (int32_t k = 0; k < n; ++k) out[k] = pcg_next(r);
    return ASTRO_OK;
}

uint32_t oracle_pcg32_next(uint64_t* st, uint64_t inc) {
    Rng r{*st, inc};
    const uint32_t v = pcg_next(r);
    *st = r.state;
    return v;
}
uint64_t oracle_cell_stream_inc(uint64_t id) { return stream_inc(id); }
uint64_t oracle_cell_rng_init(uint64_t seed, uint64_t id) { return cell_rng_init(seed, id); }
uint64_t oracle_pcg_split(uint64_t p, uint64_t d) { return pcg_split(p, d); }
double oracle_uniform01d(uint64_t* st, uint64_t inc) {
    Rng r{*st, inc};
    const double v = uniform01d(r);
    *st = r.state;
    return v;
}
void oracle_gaussian_pair(uint64_t* st, uint64_t inc, double* z0, double* z1) {
    Rng r{*st, inc};
    gaussian_pair(r, *z0, *z1);
    *st = r.state;
}
void oracle_division_axis(uint64_t did, double* ax, double* ay, double* az) { division_axis(did, *ax, *ay, *az); }
double oracle_tau_drag(double t) { return tau_drag(t); }
int32_t oracle_tau_overlaps_cell(const double a[3], const double b[3]) {
    return tau_overlaps_cell(V3{a[0], a[1], a[2]}, V3{b[0], b[1], b[2]}) ? 1 : 0;
}
int32_t oracle_tau_should_tumble(float signal, float* ema, float run, double dt) {
    return tau_should_tumble(signal, *ema, run, dt) ? 1 : 0;
}
double oracle_tau_n2_hazard(double n2, double tol) { return tau_n2_hazard(n2, tol); }
double oracle_tau_n2_death_prob(double hazard, double dt_bio) { return tau_n2_death_prob(hazard, dt_bio); }
float oracle_tau_daughter_tolerance(float parent, double mutation) { return tau_daughter_tolerance(parent, mutation); }
double oracle_ou_position_shape(double x) { return ou_position_shape(x); }
void oracle_ou_coefficients(double m, double g, double t, double dt, double o[6]) {
    const Ou c = ou_coefficients(m, g, t, dt);
    o[0] = c.e1; o[1] = c.one_minus_e1; o[2] = c.inv_beta;
    o[3] = c.sigma_rr; o[4] = c.sigma_vv; o[5] = c.sigma_rv;
}
void oracle_ou_advance_axis(double* pos, double* vel, double force, double gamma, const double c6[6],
                            double dt, int32_t noise, double z1, double z2) {
    const Ou c{c6[0], c6[1], c6[2], c6[3], c6[4], c6[5]};
    ou_advance_axis(*pos, *vel, force, gamma, c, dt, noise != 0, z1, z2);
}
double oracle_water_viscosity(double t) { return water_viscosity(t); }
double oracle_disc_overlap_fraction(double p, double a) { return disc_overlap_fraction(p, a); }
double oracle_shadow_fraction(const double s[3], const double o[3], const double l[3]) {
    return shadow_fraction(V3{s[0], s[1], s[2]}, V3{o[0], o[1], o[2]}, V3{l[0], l[1], l[2]});
}
double oracle_lumped_exchange_energy(double tc, double tg, double g, double c, double dt) {
    return lumped_exchange_energy(tc, tg, g, c, dt);
}
float oracle_taxis_ema_update(float ema, double s, double dt) { return taxis_ema_update(ema, s, dt); }
double oracle_taxis_tumble_angle(uint64_t* st, uint64_t inc) {
    Rng r{*st, inc};
    const double v = taxis_tumble_angle(r);
    *st = r.state;
    return v;
}
int64_t oracle_to_fixed(double v, double s) { return to_fixed(v, s); }
uint64_t oracle_fnv1a64(const void* d, size_t n, uint64_t h) { return fnv1a64(d, n, h); }

}  // extern "C"
