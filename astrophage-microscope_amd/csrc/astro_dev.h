// astro_dev.h -- device-side data layout and physics of the MI355X hot path.
//
// Layout in HBM (one world):
//   * cell SoA, one 256 B-aligned carve per array, in the reference's
//     contract order (contracts/cell_store_v1.h:42-74), stored physically in
//     spatial-hash bucket order; each entry carries its reference slot in
//     `lslot` (NO_SLOT for a tombstone). Position and velocity are `P`
//     (double for the contract layout, float for the hot layout); every other
//     array keeps its contract type.
//   * a second SoA of the same shape, the target of the periodic re-sort.
//   * the spatial hash as bucket-sorted neighbour records (position + slot +
//     alive bit), a dense bucket prefix `bstart[nb + 2]`, and the
//     physical <-> sorted-position maps, rebuilt every tick.
//   * per field: value, ping-pong scratch, int64 fixed-point accumulator.
//   * a small Counters block: the population count, next id and the per-tick
//     tallies live on the device, so a tick needs no host synchronisation.
//
// Energies, heat and nutrient budgets and the stats are double precision in
// both layouts. In the hot layout the integrator's drag, OU coefficients and
// Gaussian draws and the contact arithmetic are float, and displacements are
// summed onto the stored coordinates in double. Compiled with
// -ffp-contract=off so expressions round as written.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "astro_canon.h"
#include "astro_hot.h"

namespace astro_dev {

// ---------------------------------------------------------------------------
// Views passed by value to kernels
//
// Physical vs logical order. The reference's slot order (its SoA index) is the
// spawn order, which is spatially random, so a kernel walking cells in slot
// order gathers field voxels and neighbours at random. Here the SoA is kept
// *physically* in spatial-hash order and every cell carries its reference
// slot in `lslot`. Everything the reference can observe -- downloads, the
// hash's vals_sorted, division slots and ids, compaction -- is computed in
// logical-slot order; only the storage order differs. Physical entries whose
// cell was compacted away are tombstones (lslot == NO_SLOT, flags == 0) until
// the next re-sort packs them out.
// ---------------------------------------------------------------------------
constexpr uint32_t NO_SLOT = 0xFFFFFFFFu;

template <class P>
struct Cells {
    uint32_t* lslot;     // reference slot of this physical entry, or NO_SLOT
    uint64_t* id;
    uint32_t* flags;
    uint8_t*  death_cause;
    P *x, *y, *z, *vx, *vy, *vz;
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
};

// Device-resident bookkeeping (src/sim/cell_store.cuh:36-59,
// src/sim/world.cuh:111-121), so no tick has to read a count back.
struct Counters {
    int32_t  count;            // reference slots [0, count) may be occupied
    int32_t  capacity;         // reference capacity; physical extent is 2x
    int32_t  nphys;            // physical entries [0, nphys) in use (incl. tombstones)
    int32_t  resort;           // 1 while this tick re-sorts the SoA
    uint32_t tick_ctr;         // ticks run, for the re-sort cadence
    uint32_t blocks_done;      // finished-block ticket of the division pass
    uint64_t next_id;
    int64_t  divisions_total;
    int64_t  deaths_total;
    int32_t  dead_slots_prev;
    int32_t  births;           // this tick (unclamped)
    int32_t  deads;            // occupied-but-dead slots this tick
    int32_t  live;             // compaction survivors
    int32_t  div_total;        // total of this tick's divide-flag scan
    int32_t  drop_total;       // total of this tick's drop-flag scan
    uint32_t error_flags;      // sticky device-side errors (see ERR_*)
    int32_t  _pad;
    int64_t  cell_steps;       // sum of the count each tick began with
    uint64_t tick_count;       // ticks begun (the host's World::tick between steps)
    uint64_t rng_tick;         // the tick in progress: counter word of Philox streams
    uint32_t tqueued[2];       // thermostat substeps by parity: a deposit waits in the accumulator
};

enum : uint32_t { ERR_SCAN = 1u, ERR_COEFF = 2u };

// One scalar field (contracts/fields_v1.h:164-174), extended to nz planes.
struct Grid {
    float* value;
    float* scratch;
    unsigned long long* deposit;
    int32_t nx, ny, nz;
    float   fdx, fdy, fdz;
    double  scale;
};

// Spatial-hash geometry (src/sim/hash.cuh:25-34).
struct HashGeo {
    int32_t nx, ny, nz, nb;
    float   inv, ox, oy, oz;
};

// A bucket-sorted neighbour record: position plus slot (bits 0-29),
// "unoccupied" (bit 30) and "alive" (bit 31).
template <class P> struct Rec;
template <> struct alignas(32) Rec<double> { double x, y, z; uint32_t meta, _pad; };
template <> struct alignas(16) Rec<float>  { float x, y, z; uint32_t meta; };

constexpr uint32_t REC_SLOT_MASK = 0x3FFFFFFFu;
constexpr uint32_t REC_UNOCC = 1u << 30;
constexpr uint32_t REC_ALIVE = 1u << 31;

// A cell's force at the layout's width: three doubles (24 bytes) in the
// contract layout, three floats (12 bytes) in the FP32 layout, whose
// integrator takes the force in float. The buffer crosses HBM about four
// times per cell per tick (sample, contact read and write, integrate).
template <class P> struct Force { P x, y, z; };

// ---------------------------------------------------------------------------
// PCG32 per-cell streams (src/core/rng.cuh:23-117)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// The uniforms and Gaussians every generator hands out, built on its 32-bit
// draws (src/core/rng.cuh:76-101): 53-bit doubles from two draws, Box-Muller
// pairs from two doubles.
template <class G>
struct Draws {
    __device__ __forceinline__ double uniform01d() {
        G& g = static_cast<G&>(*this);
        const uint64_t hi = static_cast<uint64_t>(g.next()) << 21;
        const uint64_t lo = static_cast<uint64_t>(g.next()) >> 11;
        return static_cast<double>((hi ^ lo) & ((1ull << 53) - 1)) * (1.0 / 9007199254740992.0);
    }
    __device__ __forceinline__ double uniform(double lo, double hi) { return lo + (hi - lo) * uniform01d(); }
    __device__ __forceinline__ void gaussian_pair(double& z0, double& z1) {
        double u1 = uniform01d();
        if (u1 < 1.0e-300) u1 = 1.0e-300;
        const double u2 = uniform01d();
        const double mag = sqrt(-2.0 * log(u1));
        const double ang = 6.283185307179586476925286766559 * u2;
        double s, c;
        sincos(ang, &s, &c);
        z0 = mag * c;
        z1 = mag * s;
    }
};

struct Pcg : Draws<Pcg> {
    uint64_t state, inc;
    __device__ __forceinline__ Pcg(uint64_t s, uint64_t i) : state(s), inc(i) {}
    __device__ __forceinline__ uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t xs = static_cast<uint32_t>(((old >> 18) ^ old) >> 27);
        const uint32_t rot = static_cast<uint32_t>(old >> 59);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): ten rounds over a 128-bit counter, each two
// 32 x 32 -> 64-bit products whose halves are crossed with the other two
// words and the key; the 64-bit key is advanced by the Weyl constants between
// rounds. The counter-based alternative to the per-cell PCG32 state.
// ---------------------------------------------------------------------------
constexpr uint32_t PHILOX_M0 = 0xD2511F53u, PHILOX_M1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W0 = 0x9E3779B9u, PHILOX_W1 = 0xBB67AE85u;

__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
        const uint32_t lo0 = PHILOX_M0 * c.x, hi0 = __umulhi(PHILOX_M0, c.x);
        const uint32_t lo1 = PHILOX_M1 * c.z, hi1 = __umulhi(PHILOX_M1, c.z);
        c = uint4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    }
    return c;
}

// Stages that draw from a cell's counter stream; each starts its own run of
// blocks, so no two stages of a tick share a counter.
enum : uint32_t { RNG_STAGE_SPAWN = ASTRO_RNG_STAGE_SPAWN, RNG_STAGE_TAXIS = ASTRO_RNG_STAGE_TAXIS,
                  RNG_STAGE_INTEGRATE = ASTRO_RNG_STAGE_INTEGRATE };

// A cell's draws within one stage of one tick: key = the seed, counter =
// (stage << 16 | block, tick, id low, id high); each block yields four draws,
// handed out in order.
struct Philox : Draws<Philox> {
    uint32_t k0, k1, tick, id_lo, id_hi, word;
    uint4 buf;
    int left;
    __device__ __forceinline__ Philox(uint64_t seed, uint32_t stage, uint64_t tick_, uint64_t id)
        : k0(static_cast<uint32_t>(seed)), k1(static_cast<uint32_t>(seed >> 32)),
          tick(static_cast<uint32_t>(tick_)), id_lo(static_cast<uint32_t>(id)),
          id_hi(static_cast<uint32_t>(id >> 32)), word(stage << 16), buf{0u, 0u, 0u, 0u}, left(0) {}
    __device__ __forceinline__ uint32_t next() {
        if (left == 0) {
            buf = philox4x32_10(uint4{word, tick, id_lo, id_hi}, k0, k1);
            ++word;
            left = 4;
        }
        const uint32_t v = left == 4 ? buf.x : (left == 3 ? buf.y : (left == 2 ? buf.z : buf.w));
        --left;
        return v;
    }
};

// Which per-cell generator a world draws from, and what a Philox stream needs.
struct RngCtx {
    uint64_t seed;          // Philox key
    const uint64_t* tick;   // device word holding the tick in progress (Counters::rng_tick)
    int philox;             // 0: the cell's PCG32 state in rng_state
};

// The Gaussian pair in float for the FP32 storage layout: the same two
// uniform01d draws (the stream stays bit-exact), the transform in float.
template <class G>
__device__ __forceinline__ void gaussian_pair_f(G& r, float& z0, float& z1) {
    float u1 = static_cast<float>(r.uniform01d());
    if (!(u1 >= 1.17549435e-38f)) u1 = 1.17549435e-38f;
    const float u2 = static_cast<float>(r.uniform01d());
    const float mag = sqrtf(-2.0f * logf(u1));
    float s, c;
    sincosf(6.28318530717958648f * u2, &s, &c);
    z0 = mag * c;
    z1 = mag * s;
}

__device__ __forceinline__ uint64_t stream_inc(uint64_t id) {
    return (splitmix64(id ^ 0xA5A5A5A5DEADBEEFull) << 1u) | 1u;
}
__device__ __forceinline__ Pcg cell_rng(uint64_t st, uint64_t id) { return Pcg{st, stream_inc(id)}; }
__device__ __forceinline__ uint64_t cell_rng_init(uint64_t seed, uint64_t id) {
    Pcg r{0ull, (splitmix64(id + 1) << 1u) | 1u};
    r.next();
    r.state += seed ^ splitmix64(id);
    r.next();
    return r.state;
}
__device__ __forceinline__ uint64_t pcg_split(uint64_t parent, uint64_t did) {
    return splitmix64(parent ^ splitmix64(did * 0x2545F4914F6CDD1Dull));
}

// ---------------------------------------------------------------------------
// Fixed point (src/core/fixed_atomic.cuh:27-44)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t to_fixed(double v, double scale) { return static_cast<int64_t>(llround(v * scale)); }
__device__ __forceinline__ double from_fixed(unsigned long long raw, double scale) {
    return static_cast<double>(static_cast<int64_t>(raw)) / scale;
}
__device__ __forceinline__ void deposit_add(unsigned long long* acc, double v, double scale) {
    atomicAdd(acc, static_cast<unsigned long long>(to_fixed(v, scale)));
}

// ---------------------------------------------------------------------------
// Physics (src/sim/integrator.cuh, thermal.cuh, emission.cuh, taxis.cuh,
// lifecycle.cuh), same expressions as the reference so they round alike.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double water_viscosity(double t) { return ASTRO_VF_A * exp(ASTRO_VF_B / (t - ASTRO_VF_C)); }
__device__ __forceinline__ double drag_coefficient(double t) { return 6.0 * ASTRO_PI * water_viscosity(t) * ASTRO_CELL_RADIUS; }
__device__ __forceinline__ double cell_mass(double biomass, double energy) {
    return biomass + energy / (ASTRO_C_LIGHT * ASTRO_C_LIGHT);
}
__device__ __forceinline__ double buoyant_weight(double mass) {
    return -(mass - ASTRO_WATER_DENSITY * ASTRO_CELL_VOLUME) * ASTRO_G_ACCEL;
}

__device__ __forceinline__ double ou_position_shape(double x) {
    if (x < 1.0e-2) {
        const double x3 = x * x * x;
        return x3 * (2.0 / 3.0 + x * (-0.5 + x * (7.0 / 30.0 + x * (-1.0 / 12.0))));
    }
    const double e = exp(-x);
    return 2.0 * x - 3.0 + 4.0 * e - e * e;
}

// The joint OU step's coefficients; noise_a / noise_b are the Cholesky
// factors ou_advance_axis derives from the sigmas, the same for every axis.
struct Ou { double one_minus_e1, inv_beta, sigma_rr, sigma_vv, sigma_rv, noise_a, noise_b; };

__device__ __forceinline__ Ou ou_coefficients(double mass, double gamma, double t, double dt) {
    const double beta = gamma / mass;
    const double x = beta * dt;
    Ou c;
    c.one_minus_e1 = -expm1(-x);
    c.inv_beta = 1.0 / beta;
    const double kT_over_m = ASTRO_K_BOLTZ * t / mass;
    const double var_vv = kT_over_m * (-expm1(-2.0 * x));
    const double var_rr = kT_over_m * c.inv_beta * c.inv_beta * ou_position_shape(x);
    c.sigma_vv = sqrt(var_vv > 0.0 ? var_vv : 0.0);
    c.sigma_rr = sqrt(var_rr > 0.0 ? var_rr : 0.0);
    c.sigma_rv = kT_over_m * c.inv_beta * c.one_minus_e1 * c.one_minus_e1;
    c.noise_a = 0.0;
    c.noise_b = 0.0;
    if (c.sigma_rr > 0.0) {
        c.noise_a = c.sigma_rv / c.sigma_rr;
        const double b2 = c.sigma_vv * c.sigma_vv - c.noise_a * c.noise_a;
        c.noise_b = sqrt(b2 > 0.0 ? b2 : 0.0);
    }
    return c;
}

// The FP32 storage layout's integrator: drag, coefficients and the noise in
// float; displacements are summed onto the position and velocity in double,
// where the float terms add exactly, so the only rounding of the result is
// the float store, as in the double path, and the wall tests see the same
// values to within the float terms' own error (about 1e-7 of a
// displacement). Small beta * dt, where the float forms would cancel, takes
// the double coefficients.
struct OuF { float one_minus_e1, inv_beta, sigma_rr, sigma_vv, noise_a, noise_b; };

__device__ __forceinline__ float drag_coefficient_f(float t) {
    const float eta = static_cast<float>(ASTRO_VF_A) * expf(static_cast<float>(ASTRO_VF_B) / (t - static_cast<float>(ASTRO_VF_C)));
    return 6.0f * static_cast<float>(ASTRO_PI) * eta * static_cast<float>(ASTRO_CELL_RADIUS);
}

__device__ __forceinline__ OuF ou_coefficients_f(float mass, float gamma, float t, float dt) {
    const float beta = gamma / mass;
    const float x = beta * dt;
    OuF c;
    if (x < 1.0f) {
        const Ou d = ou_coefficients(static_cast<double>(mass), static_cast<double>(gamma), static_cast<double>(t),
                                     static_cast<double>(dt));
        c.one_minus_e1 = static_cast<float>(d.one_minus_e1);
        c.inv_beta = static_cast<float>(d.inv_beta);
        c.sigma_rr = static_cast<float>(d.sigma_rr);
        c.sigma_vv = static_cast<float>(d.sigma_vv);
        c.noise_a = static_cast<float>(d.noise_a);
        c.noise_b = static_cast<float>(d.noise_b);
        return c;
    }
    c.one_minus_e1 = -expm1f(-x);
    c.inv_beta = 1.0f / beta;
    const float kT_over_m = static_cast<float>(ASTRO_K_BOLTZ) * t / mass;
    const float var_vv = kT_over_m * (-expm1f(-2.0f * x));
    const float e = expf(-x);
    const float shape = 2.0f * x - 3.0f + 4.0f * e - e * e;
    const float var_rr = kT_over_m * c.inv_beta * c.inv_beta * shape;
    c.sigma_vv = sqrtf(var_vv > 0.0f ? var_vv : 0.0f);
    c.sigma_rr = sqrtf(var_rr > 0.0f ? var_rr : 0.0f);
    const float sigma_rv = kT_over_m * c.inv_beta * c.one_minus_e1 * c.one_minus_e1;
    c.noise_a = 0.0f;
    c.noise_b = 0.0f;
    if (c.sigma_rr > 0.0f) {
        c.noise_a = sigma_rv / c.sigma_rr;
        const float b2 = c.sigma_vv * c.sigma_vv - c.noise_a * c.noise_a;
        c.noise_b = sqrtf(b2 > 0.0f ? b2 : 0.0f);
    }
    return c;
}

__device__ __forceinline__ void ou_advance_axis_f(double& pos, double& vel, float force, float gamma, const OuF& c,
                                                  float dt, bool noise, float z1, float z2) {
    const float v = static_cast<float>(vel);
    const float v_term = force / gamma;
    const float dpos = v_term * dt + (v - v_term) * c.one_minus_e1 * c.inv_beta;
    const float dvel = (v_term - v) * c.one_minus_e1;
    double npos = pos + static_cast<double>(dpos);
    double nvel = vel + static_cast<double>(dvel);
    if (noise && c.sigma_rr > 0.0f) {
        npos += static_cast<double>(c.sigma_rr * z1);
        nvel += static_cast<double>(c.noise_a * z1 + c.noise_b * z2);
    } else if (noise) {
        nvel += static_cast<double>(c.sigma_vv * z2);
    }
    pos = npos;
    vel = nvel;
}

__device__ __forceinline__ void ou_advance_axis(double& pos, double& vel, double force, double gamma,
                                                const Ou& c, double dt, bool noise, double z1, double z2) {
    const double v_term = force / gamma;
    const double dpos = v_term * dt + (vel - v_term) * c.one_minus_e1 * c.inv_beta;
    const double dvel = (v_term - vel) * c.one_minus_e1;
    double npos = pos + dpos;
    double nvel = vel + dvel;
    if (noise && c.sigma_rr > 0.0) {
        npos += c.sigma_rr * z1;
        nvel += c.noise_a * z1 + c.noise_b * z2;
    } else if (noise) {
        nvel += c.sigma_vv * z2;
    }
    pos = npos;
    vel = nvel;
}

__device__ __forceinline__ bool apply_boundary_axis(double& p, double& v, double half, double radius, int mode) {
    const double lo = -half + radius;
    const double hi = half - radius;
    if (p >= lo && p <= hi) return true;
    if (mode == ASTRO_BOUNDARY_PERIODIC) {
        const double span = 2.0 * half;
        while (p < -half) p += span;
        while (p > half) p -= span;
        return true;
    }
    if (mode == ASTRO_BOUNDARY_ABSORBING) return false;
    p = (p < lo) ? lo : hi;
    v = 0.0;
    return true;
}

// division_axis (src/sim/lifecycle.cuh:87-99): a unit vector hashed from the
// daughter id, so placing the pair consumes no draw.
__device__ __forceinline__ void division_axis(uint64_t did, double& ax, double& ay, double& az) {
    const uint64_t h = splitmix64(splitmix64(did));
    const double u = static_cast<double>(h >> 33) * (1.0 / 2147483648.0) - 1.0;
    const double phi = static_cast<double>(h & 0xFFFFFFFFull) * (ASTRO_TWO_PI / 4294967296.0);
    const double sq = sqrt(u * u < 1.0 ? 1.0 - u * u : 0.0);
    ax = sq * cos(phi);
    ay = sq * sin(phi);
    az = u;
}

__device__ __forceinline__ double disc_overlap_fraction(double p, double a) {
    if (p <= 0.0) return 1.0;
    if (p >= 2.0 * a) return 0.0;
    const double x = p / (2.0 * a);
    return (2.0 / ASTRO_PI) * (acos(x) - x * sqrt(1.0 - x * x));
}

__device__ __forceinline__ double column_transmittance(double blocked, double face) {
    if (face <= 0.0) return 1.0;
    const double t = 1.0 - blocked / face;
    return t > 0.0 ? t : 0.0;
}

__device__ __forceinline__ double co2_uptake_rate(double c) {
    if (c <= 0.0) return 0.0;
    return ASTRO_LIFE_CO2_UPTAKE_MAX * c / (c + ASTRO_LIFE_CO2_HALF_SATURATION);
}

// ---------------------------------------------------------------------------
// Field access (src/fields/grid.cuh:109-141), trilinear/voxel when nz > 1.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int clampi(int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); }

__device__ __forceinline__ float grid_sample(const Grid& g, float x, float y, float z) {
    const float half_x = 0.5f * g.nx * g.fdx;
    const float half_y = 0.5f * g.ny * g.fdy;
    const float gx = (x + half_x) / g.fdx - 0.5f;
    const float gy = (y + half_y) / g.fdy - 0.5f;
    const int x0 = static_cast<int>(floorf(gx));
    const int y0 = static_cast<int>(floorf(gy));
    const float tx = gx - static_cast<float>(x0);
    const float ty = gy - static_cast<float>(y0);
    const int j0 = clampi(y0, g.ny), j1 = clampi(y0 + 1, g.ny);
    const size_t pl = static_cast<size_t>(g.nx) * g.ny;
    // Values at (i0, j) and (i1, j) of one row as one 8-byte load: the pair
    // (b, b + 1) with b = clamp(x0, 0, nx - 2) holds both, also at the clamped
    // edges (x0 < 0: both are element b; x0 > nx - 2: both are b + 1).
    const int xb = x0 < 0 ? 0 : (x0 > g.nx - 2 ? g.nx - 2 : x0);
    const bool lo_edge = x0 < 0, hi_edge = x0 > g.nx - 2;
    auto pair = [&](const float* row, float& a, float& b) {
        float2 q;
        __builtin_memcpy(&q, row + xb, sizeof(q));
        a = hi_edge ? q.y : q.x;
        b = lo_edge ? q.x : q.y;
    };
    auto plane = [&](int k) -> float {
        const float* v = g.value + static_cast<size_t>(k) * pl;
        float v00, v10, v01, v11;
        pair(v + static_cast<size_t>(j0) * g.nx, v00, v10);
        pair(v + static_cast<size_t>(j1) * g.nx, v01, v11);
        return (1.0f - tx) * (1.0f - ty) * v00 + tx * (1.0f - ty) * v10 + (1.0f - tx) * ty * v01 + tx * ty * v11;
    };
    if (g.nz == 1) return plane(0);
    const float half_z = 0.5f * g.nz * g.fdz;
    const float gz = (z + half_z) / g.fdz - 0.5f;
    const int z0 = static_cast<int>(floorf(gz));
    const float tz = gz - static_cast<float>(z0);
    const float b0 = plane(clampi(z0, g.nz));
    const float b1 = plane(clampi(z0 + 1, g.nz));
    return (1.0f - tz) * b0 + tz * b1;
}

__device__ __forceinline__ uint32_t grid_nearest(const Grid& g, float x, float y, float z) {
    const float half_x = 0.5f * g.nx * g.fdx;
    const float half_y = 0.5f * g.ny * g.fdy;
    const int i = clampi(static_cast<int>((x + half_x) / g.fdx), g.nx);
    const int j = clampi(static_cast<int>((y + half_y) / g.fdy), g.ny);
    int k = 0;
    if (g.nz > 1) {
        const float half_z = 0.5f * g.nz * g.fdz;
        k = clampi(static_cast<int>((z + half_z) / g.fdz), g.nz);
    }
    return static_cast<uint32_t>((static_cast<size_t>(k) * g.ny + j) * g.nx + i);
}

// ---------------------------------------------------------------------------
// Spatial hash (src/sim/hash.cuh:59-78)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void hash_coords(const HashGeo& h, double x, double y, double z,
                                            int& gx, int& gy, int& gz) {
    auto cl = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
    gx = cl(static_cast<int32_t>((x - static_cast<double>(h.ox)) * static_cast<double>(h.inv)), h.nx - 1);
    gy = cl(static_cast<int32_t>((y - static_cast<double>(h.oy)) * static_cast<double>(h.inv)), h.ny - 1);
    gz = cl(static_cast<int32_t>((z - static_cast<double>(h.oz)) * static_cast<double>(h.inv)), h.nz - 1);
}
__device__ __forceinline__ uint32_t hash_key(const HashGeo& h, double x, double y, double z) {
    int gx, gy, gz;
    hash_coords(h, x, y, z, gx, gy, gz);
    return static_cast<uint32_t>((gz * h.ny + gy) * h.nx + gx);
}

// Copies one cell's whole state (every SoA array, incl. lslot).
template <class P>
__device__ __forceinline__ void copy_cell(const Cells<P>& s, int i, const Cells<P>& d, int k) {
    d.lslot[k] = s.lslot[i];
    d.id[k] = s.id[i]; d.flags[k] = s.flags[i]; d.death_cause[k] = s.death_cause[i];
    d.x[k] = s.x[i]; d.y[k] = s.y[i]; d.z[k] = s.z[i];
    d.vx[k] = s.vx[i]; d.vy[k] = s.vy[i]; d.vz[k] = s.vz[i];
    d.energy[k] = s.energy[i]; d.temp_cell[k] = s.temp_cell[i]; d.emit_power[k] = s.emit_power[i];
    d.dir_x[k] = s.dir_x[i]; d.dir_y[k] = s.dir_y[i]; d.dir_z[k] = s.dir_z[i];
    d.biomass[k] = s.biomass[i]; d.co2_held[k] = s.co2_held[i]; d.age_s[k] = s.age_s[i];
    d.rng_state[k] = s.rng_state[i]; d.taxis_memory[k] = s.taxis_memory[i]; d.run_timer[k] = s.run_timer[i];
    d.t_local[k] = s.t_local[i]; d.irradiance[k] = s.irradiance[i]; d.co2_local[k] = s.co2_local[i];
    d.n2_local[k] = s.n2_local[i];
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
constexpr int BLOCK = 256;
constexpr int NUM_XCD = 8;
// Per-cell kernels loop grid-stride over a grid of at most MAX_BLOCKS (8
// resident blocks per CU), so a grid sized for the capacity costs nothing
// when the population is smaller and conditional passes exit cheaply.
constexpr int MAX_BLOCKS = 2048;
__host__ __device__ __forceinline__ int stride_grid(int n) {
    const int b = (n + BLOCK - 1) / BLOCK;
    return b < 1 ? 1 : (b < MAX_BLOCKS ? b : MAX_BLOCKS);
}

// Remaps a linear block id so each XCD (blocks are dealt to the 8 XCDs round
// robin) gets one contiguous run of tiles, which keeps neighbouring tiles in
// the same L2. Launch `blocks_for_xcd(n)` blocks; ids >= n come back >= n.
__host__ __device__ __forceinline__ int blocks_for_xcd(int n) { return ((n + NUM_XCD - 1) / NUM_XCD) * NUM_XCD; }
__device__ __forceinline__ int xcd_tile(int b, int nlaunched) {
    const int per = nlaunched / NUM_XCD;
    return (b % NUM_XCD) * per + b / NUM_XCD;
}

// One block per BLOCK-entry tile of n entries, a multiple of NUM_XCD. With
// more blocks than fit at once, blocks start as slots free up and a CU's
// waves drift out of step, so one wave's memory waits overlap another's
// arithmetic (passes that walk tiles with xcd_tiles work with any grid).
__host__ __device__ __forceinline__ int tile_grid(int n) { return blocks_for_xcd(n < 1 ? 1 : (n + BLOCK - 1) / BLOCK); }

// Grid size of a per-cell pass over up to n entries: at most MAX_BLOCKS, a
// multiple of NUM_XCD.
__host__ __device__ __forceinline__ int cell_grid(int n) { return blocks_for_xcd(stride_grid(n)); }

// A kernel body's own block index and block count: the launch's for a
// single world, the chamber's share of it for a batched launch.
__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = min(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
    return v;
}
// Orders one wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct VGrid {
    int block, grid;
};

// Runs body(i, valid) for i over [0, n) in BLOCK-wide tiles. The tiles are
// split into NUM_XCD contiguous chunks and the blocks of XCD x walk chunk x,
// so the cells one L2 serves are spatial neighbours (cells are stored in
// spatial-hash order) and share field voxels and neighbour records. Whole
// blocks iterate together (`valid` is false past n), so the body may use
// wave-wide operations. A grid that is not a multiple of NUM_XCD (a small
// chamber of a batched launch) walks the tiles in plain grid-stride order.
template <class Body>
__device__ __forceinline__ void xcd_tiles(VGrid vg, int n, Body&& body) {
    const int ntiles = (n + BLOCK - 1) / BLOCK;
    if (vg.grid % NUM_XCD != 0) {
        for (int t = vg.block; t < ntiles; t += vg.grid) {
            const int i = t * BLOCK + static_cast<int>(threadIdx.x);
            body(i, i < n);
        }
        return;
    }
    const int x = vg.block % NUM_XCD;
    const int nb = vg.grid / NUM_XCD;
    const int per = (ntiles + NUM_XCD - 1) / NUM_XCD;
    const int t1 = min(ntiles, (x + 1) * per);
    for (int t = x * per + vg.block / NUM_XCD; t < t1; t += nb) {
        const int i = t * BLOCK + static_cast<int>(threadIdx.x);
        body(i, i < n);
    }
}

}  // namespace astro_dev
