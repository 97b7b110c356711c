// cell_ops.h -- per-cell stage bodies shared by the stand-alone stage kernels
// and the fused per-tick pass. Each touches only cell i's own state.
#pragma once

#include "astro_dev.h"

namespace astro_dev {

__device__ __forceinline__ void any_perpendicular(double hx, double hy, double hz, double& ux, double& uy, double& uz) {
    const double ax = fabs(hx), ay = fabs(hy), az = fabs(hz);
    double bx = 0.0, by = 0.0, bz = 0.0;
    if (ax <= ay && ax <= az) bx = 1.0;
    else if (ay <= az) by = 1.0;
    else bz = 1.0;
    double cx = hy * bz - hz * by, cy = hz * bx - hx * bz, cz = hx * by - hy * bx;
    const double l = sqrt(cx * cx + cy * cy + cz * cz);
    if (l > 0.0) {
        const double inv = 1.0 / l;
        ux = cx * inv; uy = cy * inv; uz = cz * inv;
    } else {
        ux = 0.0; uy = 0.0; uz = 1.0;
    }
}

// feed_kernel's absorption half (src/sim/emission.cu:107-114) followed by
// taxis_kernel (src/sim/taxis.cu:13-79) for one cell: both touch only the
// cell's own irradiance, energy and emission state, in that order.
// `directional` is the cell's shadowed far-field irradiance (FEED only).
template <class P, bool FEED, bool TAXIS>
__device__ __forceinline__ void feed_taxis_cell(const Cells<P>& c, int i, uint32_t f, double directional,
                                                float ambient, double dt, double max_power, const RngCtx& rng) {
    if (!(f & ASTRO_FLAG_OCCUPIED) || !(f & ASTRO_FLAG_ALIVE)) return;
    double energy = c.energy[i];
    float irr_f = 0.0f;
    if constexpr (FEED) {
        irr_f = static_cast<float>(directional);
        c.irradiance[i] = irr_f;
        const double gained = (directional * ASTRO_CELL_CROSS_SECTION +
                               static_cast<double>(ambient) * ASTRO_CELL_SURFACE_AREA) * dt;
        double e = energy + gained;
        if (e > ASTRO_CELL_ENERGY_MAX) e = ASTRO_CELL_ENERGY_MAX;
        energy = e;
        c.energy[i] = e;
    }
    if constexpr (TAXIS) {
        if (!(f & ASTRO_FLAG_AWAKE)) {
            c.emit_power[i] = 0.0f;
            return;
        }
        if constexpr (!FEED) irr_f = c.irradiance[i];
        const double irr = static_cast<double>(irr_f);
        const double co2 = static_cast<double>(c.co2_local[i]);
        const double charge = energy / ASTRO_CELL_ENERGY_MAX;
        // taxis_select_state (src/sim/taxis.cuh:55-62)
        const bool dark = irr < ASTRO_TAXIS_DARK_THRESHOLD;
        const bool co2_here = co2 > 0.0;
        int state = 0;
        if (dark && !co2_here) state = 0;
        else if (!dark && charge < ASTRO_TAXIS_SEEK_FEED_BELOW) state = 1;
        else if (charge > ASTRO_TAXIS_SEEK_BREED_ABOVE && co2_here) state = 2;
        if (state == 0) {
            c.emit_power[i] = 0.0f;
            c.run_timer[i] = 0.0f;
            c.taxis_memory[i] = 0.0f;
            return;
        }
        const double signal = state == 1 ? irr : co2;
        const float mem = c.taxis_memory[i];
        const double delta = signal - static_cast<double>(mem);
        const double alpha = 1.0 - exp(-dt / ASTRO_TAXIS_MEMORY_TIME);
        c.taxis_memory[i] = static_cast<float>(static_cast<double>(mem) + alpha * (signal - static_cast<double>(mem)));
        double hx = -static_cast<double>(c.dir_x[i]);
        double hy = -static_cast<double>(c.dir_y[i]);
        double hz = -static_cast<double>(c.dir_z[i]);
        double run = static_cast<double>(c.run_timer[i]) + dt;
        if (delta <= 0.0 || run >= ASTRO_TAXIS_RUN_MAX) {
            // taxis_tumble (src/sim/taxis.cuh:108-118)
            double u_theta, u_azimuth;
            if (rng.philox) {
                Philox r(rng.seed, RNG_STAGE_TAXIS, *rng.tick, c.id[i]);
                u_theta = r.uniform01d();
                u_azimuth = r.uniform01d();
            } else {
                Pcg r = cell_rng(c.rng_state[i], c.id[i]);
                u_theta = r.uniform01d();
                u_azimuth = r.uniform01d();
                c.rng_state[i] = r.state;
            }
            double theta = -ASTRO_TAXIS_TUMBLE_ANGLE_MEAN * log(1.0 - u_theta);
            if (theta > ASTRO_PI) theta = ASTRO_PI;
            const double azimuth = ASTRO_TWO_PI * u_azimuth;
            double l = sqrt(hx * hx + hy * hy + hz * hz);
            if (l > 0.0) {
                const double inv = 1.0 / l;
                hx *= inv; hy *= inv; hz *= inv;
            } else {
                hx = 0.0; hy = 0.0; hz = 1.0;
            }
            double ux, uy, uz;
            any_perpendicular(hx, hy, hz, ux, uy, uz);
            const double wx = hy * uz - hz * uy, wy = hz * ux - hx * uz, wz = hx * uy - hy * ux;
            const double ca = cos(azimuth), sa = sin(azimuth);
            const double px = ux * ca + wx * sa, py = uy * ca + wy * sa, pz = uz * ca + wz * sa;
            const double ct = cos(theta), st = sin(theta);
            double nx = hx * ct + px * st, ny = hy * ct + py * st, nz = hz * ct + pz * st;
            l = sqrt(nx * nx + ny * ny + nz * nz);
            if (l > 0.0) {
                const double inv = 1.0 / l;
                nx *= inv; ny *= inv; nz *= inv;
            } else {
                nx = 0.0; ny = 0.0; nz = 1.0;
            }
            hx = nx; hy = ny; hz = nz;
            run = 0.0;
        }
        c.run_timer[i] = static_cast<float>(run);
        // taxis_emit_dir = -normalize(heading)
        double l = sqrt(hx * hx + hy * hy + hz * hz);
        double ex, ey, ez;
        if (l > 0.0) {
            const double inv = 1.0 / l;
            ex = hx * inv; ey = hy * inv; ez = hz * inv;
        } else {
            ex = 0.0; ey = 0.0; ez = 1.0;
        }
        c.dir_x[i] = static_cast<float>(-ex);
        c.dir_y[i] = static_cast<float>(-ey);
        c.dir_z[i] = static_cast<float>(-ez);
        double p = 0.0;
        if (energy > 0.0) {
            const double avail = dt > 0.0 ? energy / dt : 0.0;
            p = max_power < avail ? max_power : avail;
        }
        c.emit_power[i] = static_cast<float>(p);
        c.energy[i] = energy - p * dt;
    }
}

// field_sample_kernel (src/sim/integrator.cu:20-30) and the body-force half
// of forces_kernel (:40-58) for one cell: samples the fields into the cell's
// t/co2/n2_local and returns buoyancy + photon recoil (zero if unoccupied).
// (x, y, z) is the cell's position, as the float the samplers take.
template <class P>
__device__ __forceinline__ Force<double> sample_force_cell(const Cells<P>& c, int i, uint32_t fl, float x, float y, float z,
                                                   const Grid& T, const Grid& C, const Grid& N, int c_on, int n_on,
                                                   int gaxis) {
    Force<double> F{0.0, 0.0, 0.0};
    if (!(fl & ASTRO_FLAG_OCCUPIED)) return F;
    c.t_local[i] = grid_sample(T, x, y, z);
    c.co2_local[i] = c_on ? grid_sample(C, x, y, z) : 0.0f;
    c.n2_local[i] = n_on ? grid_sample(N, x, y, z) : 0.0f;
    const double mass = cell_mass(c.biomass[i], c.energy[i]);
    const bool alive = (fl & ASTRO_FLAG_ALIVE) != 0u;
    const double emit = alive ? static_cast<double>(c.emit_power[i]) : 0.0;
    const double w = buoyant_weight(mass);
    if (gaxis == ASTRO_GRAVITY_Y) F.y += w;
    else F.z += w;
    const double thrust = emit / ASTRO_C_LIGHT;
    if (thrust != 0.0) {
        F.x -= static_cast<double>(c.dir_x[i]) * thrust;
        F.y -= static_cast<double>(c.dir_y[i]) * thrust;
        F.z -= static_cast<double>(c.dir_z[i]) * thrust;
    }
    return F;
}

}  // namespace astro_dev
