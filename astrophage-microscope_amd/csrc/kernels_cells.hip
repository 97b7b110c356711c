// kernels_cells.hip -- per-cell stages: spawn, thermostat, feeding + taxis,
// field sample + body force, the OU integrator, lifecycle and compaction, and
// the per-cell telemetry reduction. One thread per slot; every per-cell array
// is read and written coalesced in slot order.
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
template <class P>
__global__ __launch_bounds__(BLOCK) void k_spawn(Cells<P> c, SpawnArgs p, int32_t base, int32_t lbase,
                                                 uint64_t first_id, uint64_t seed) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= p.count) return;
    const int32_t slot = base + i;
    c.lslot[slot] = static_cast<uint32_t>(lbase + i);
    const uint64_t id = first_id + static_cast<uint64_t>(i);
    Pcg r = cell_rng(cell_rng_init(seed, id), id);
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
    c.rng_state[slot] = r.state;
    c.taxis_memory[slot] = 0.0f;
    c.run_timer[slot] = 0.0f;
    c.t_local[slot] = tc;
    c.irradiance[slot] = 0.0f;
    c.co2_local[slot] = 0.0f;
    c.n2_local[slot] = 0.0f;
}

// ---------------------------------------------------------------------------
// thermal_kernel (src/sim/thermal.cu:19-71). A starving cell also clears the
// alive bit of its sorted neighbour record, so the feed walk later this tick
// sees the same flags the reference's feed_kernel reads.
// ---------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(BLOCK) void k_thermal(Cells<P> c, const Counters* ctr, Grid T, double dt_sub,
                                                   double cap, Rec<P>* rec, const int32_t* spos, int update_rec) {
    const int n = ctr->nphys;
    const auto body = [&](const int i) {
    const uint32_t fl = c.flags[i];
    if (!(fl & ASTRO_FLAG_OCCUPIED)) return;
    const uint32_t idx = grid_nearest(T, static_cast<float>(c.x[i]), static_cast<float>(c.y[i]),
                                      static_cast<float>(c.z[i]));
    const float t_grid = T.value[idx];
    bool awake = (fl & ASTRO_FLAG_AWAKE) != 0u;
    uint32_t nf = fl;
    if (!awake && static_cast<double>(t_grid) >= ASTRO_CELL_TEMP_SETPOINT) {
        awake = true;
        nf = fl | ASTRO_FLAG_AWAKE;
        c.flags[i] = nf;
    }
    const double t_cell = awake ? ASTRO_CELL_TEMP_SETPOINT : static_cast<double>(t_grid);
    c.temp_cell[i] = static_cast<float>(t_cell);
    if (!awake || !(fl & ASTRO_FLAG_ALIVE)) return;
    const double e = lumped_exchange_energy(t_cell, static_cast<double>(t_grid), ASTRO_CONDUCTION_COEFF, cap, dt_sub);
    const double e0 = c.energy[i];
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
    deposit_add(&T.deposit[idx], moved / cap, T.scale);
    if (energy <= 0.0) {
        c.flags[i] = nf & ~ASTRO_FLAG_ALIVE;
        c.death_cause[i] = ASTRO_DEATH_STARVED;
        c.emit_power[i] = 0.0f;
        if (update_rec) rec[spos[i]].meta &= ~REC_ALIVE;
    }
    };
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) body(i);
}

// Folds a field's fixed-point accumulator into its values at the cells that
// may have deposited (flush_kernel, src/fields/diffuse.cu:51-60, made sparse).
// Every depositor exchanges its voxel's accumulator with zero; exactly one of
// them receives the total and applies it, so the voxel sees the same single
// `value += float(from_fixed(total))` as a dense flush. AWAKE marks every
// thermostat depositor (thermal.cu:130), dead or alive afterwards.
template <class P>
__global__ __launch_bounds__(BLOCK) void k_apply_awake(Cells<P> c, const Counters* ctr, Grid g) {
    const int n = ctr->nphys;
    const auto body = [&](const int i) {
    const uint32_t fl = c.flags[i];
    if (!(fl & ASTRO_FLAG_OCCUPIED) || !(fl & ASTRO_FLAG_AWAKE)) return;
    const uint32_t idx = grid_nearest(g, static_cast<float>(c.x[i]), static_cast<float>(c.y[i]),
                                      static_cast<float>(c.z[i]));
    const unsigned long long raw = atomicExch(&g.deposit[idx], 0ull);
    if (raw != 0ull) g.value[idx] += static_cast<float>(from_fixed(raw, g.scale));
    };
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) body(i);
}

// ---------------------------------------------------------------------------
// feed_kernel's absorption half (src/sim/emission.cu:107-114) fused with
// taxis_kernel (src/sim/taxis.cu:13-79): both touch only the cell's own
// irradiance, energy and emission state, in that order.
// ---------------------------------------------------------------------------
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

template <class P, bool FEED, bool TAXIS>
__global__ __launch_bounds__(BLOCK) void k_feed_taxis(Cells<P> c, const Counters* ctr, const int32_t* spos,
                                                      const double* dirs, float ambient, double dt, double max_power) {
    const int n = ctr->nphys;
    const auto body = [&](const int i) {
    const uint32_t f = c.flags[i];
    if (!(f & ASTRO_FLAG_OCCUPIED) || !(f & ASTRO_FLAG_ALIVE)) return;
    double energy = c.energy[i];
    float irr_f = 0.0f;
    if constexpr (FEED) {
        const double directional = dirs[spos[i]];
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
            Pcg r = cell_rng(c.rng_state[i], c.id[i]);
            double theta = -ASTRO_TAXIS_TUMBLE_ANGLE_MEAN * log(1.0 - r.uniform01d());
            if (theta > ASTRO_PI) theta = ASTRO_PI;
            const double azimuth = ASTRO_TWO_PI * r.uniform01d();
            c.rng_state[i] = r.state;
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
    };
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) body(i);
}

// ---------------------------------------------------------------------------
// field_sample_kernel (src/sim/integrator.cu:20-30) + the body-force half of
// forces_kernel (:40-58): buoyancy and photon recoil, written to the cell's
// sorted position for the contact walk to continue.
// ---------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(BLOCK) void k_sample_base(Cells<P> c, const Counters* ctr, Grid T, Grid C, Grid N,
                                                       int c_on, int n_on, int gaxis, const int32_t* spos, Force* fs) {
    const int n = ctr->nphys;
    const auto body = [&](const int i) {
    if (c.lslot[i] == NO_SLOT) return;
    const uint32_t fl = c.flags[i];
    Force F{0.0, 0.0, 0.0, 0.0};
    if (fl & ASTRO_FLAG_OCCUPIED) {
        const float x = static_cast<float>(c.x[i]);
        const float y = static_cast<float>(c.y[i]);
        const float z = static_cast<float>(c.z[i]);
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
    }
    fs[spos[i]] = F;
    };
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) body(i);
}

// ---------------------------------------------------------------------------
// integrate_kernel (src/sim/integrator.cu:79-150).
// ---------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(BLOCK) void k_integrate(Cells<P> c, const Counters* ctr, const int32_t* spos,
                                                     const Force* fs, TickParams t) {
    const int n = ctr->nphys;
    const auto body = [&](const int i) {
    const uint32_t fl = c.flags[i];
    if (!(fl & ASTRO_FLAG_OCCUPIED)) return;
    const double dt = t.dt;
    const double t_local = static_cast<double>(c.t_local[i]);
    const double mass = cell_mass(c.biomass[i], c.energy[i]);
    const bool awake = (fl & ASTRO_FLAG_AWAKE) != 0u;
    const double t_drag = awake ? ASTRO_CELL_TEMP_SETPOINT : t_local;
    double gamma = drag_coefficient(t_drag);
    const Force F = fs[spos[i]];
    double px = static_cast<double>(c.x[i]), py = static_cast<double>(c.y[i]), pz = static_cast<double>(c.z[i]);
    double vx = static_cast<double>(c.vx[i]), vy = static_cast<double>(c.vy[i]), vz = static_cast<double>(c.vz[i]);
    if (fl & ASTRO_FLAG_STUCK) gamma *= ASTRO_WALL_STUCK_DRAG_MULT;
    const uint64_t id = c.id[i];
    Pcg r = cell_rng(c.rng_state[i], id);
    const Ou oc = ou_coefficients(mass, gamma, t_drag, dt);
    double z1, z2, z3, z4, z5, z6;
    r.gaussian_pair(z1, z2);
    r.gaussian_pair(z3, z4);
    r.gaussian_pair(z5, z6);
    const bool noise = t.thermal_noise != 0;
    ou_advance_axis(px, vx, F.x, gamma, oc, dt, noise, z1, z2);
    ou_advance_axis(py, vy, F.y, gamma, oc, dt, noise, z3, z4);
    ou_advance_axis(pz, vz, F.z, gamma, oc, dt, noise, z5, z6);
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
            // stored (integrator.cu:124-127); r.state is that state.
            const bool sticks = r.uniform01d() < ASTRO_WALL_STICKINESS;
            if (sticks) nf |= ASTRO_FLAG_STUCK;
        } else if (fl & ASTRO_FLAG_STUCK) {
            const double away = (pz < 0.0) ? F.z : -F.z;
            if (away > ASTRO_WALL_RELEASE_FORCE) nf &= ~ASTRO_FLAG_STUCK;
        }
    }
    c.rng_state[i] = r.state;
    if (!kept) {
        nf &= ~ASTRO_FLAG_ALIVE;
        c.death_cause[i] = ASTRO_DEATH_CULLED;
        c.emit_power[i] = 0.0f;
    }
    c.flags[i] = nf;
    c.x[i] = static_cast<P>(px);
    c.y[i] = static_cast<P>(py);
    c.z[i] = static_cast<P>(pz);
    c.vx[i] = static_cast<P>(vx);
    c.vy[i] = static_cast<P>(vy);
    c.vz[i] = static_cast<P>(vz);
    c.age_s[i] += static_cast<float>(dt);
    };
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) body(i);
}

// ---------------------------------------------------------------------------
// lifecycle_step (src/sim/lifecycle.cu:208-294), host syncs replaced by
// device counters.
// ---------------------------------------------------------------------------
constexpr uint32_t NO_VOXEL = 0xFFFFFFFFu;

// demand_kernel (lifecycle.cu:41-52). Also records the voxel each cell booked
// against (NO_VOXEL if none), for the uptake and the sparse flush.
template <class P>
__global__ __launch_bounds__(BLOCK) void k_demand(Cells<P> c, const Counters* ctr, Grid C, double dt_bio, double gv,
                                                  unsigned long long* demand, uint32_t* vox) {
    const int n = ctr->nphys;
    const auto body = [&](const int i) {
    vox[i] = NO_VOXEL;
    const uint32_t f = c.flags[i];
    if (!(f & ASTRO_FLAG_OCCUPIED) || !(f & ASTRO_FLAG_ALIVE)) return;
    const uint32_t idx = grid_nearest(C, static_cast<float>(c.x[i]), static_cast<float>(c.y[i]),
                                      static_cast<float>(c.z[i]));
    const double want = co2_uptake_rate(static_cast<double>(C.value[idx])) * dt_bio;
    if (want <= 0.0) return;
    deposit_add(&demand[idx], want / gv, C.scale);
    vox[i] = idx;
    };
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) body(i);
}

// uptake_kernel (lifecycle.cu:62-86).
template <class P>
__global__ __launch_bounds__(BLOCK) void k_uptake(Cells<P> c, const Counters* ctr, Grid C, double dt_bio, double gv,
                                                  const unsigned long long* demand, const uint32_t* vox) {
    const int n = ctr->nphys;
    const auto body = [&](const int i) {
    const uint32_t idx = vox[i];
    if (idx == NO_VOXEL) return;
    const double available = static_cast<double>(C.value[idx]);
    const double want = co2_uptake_rate(available) * dt_bio;
    if (want <= 0.0) return;
    const double asked = from_fixed(demand[idx], C.scale);
    double taken = want;
    if (asked > available && asked > 0.0) taken = want * (available / asked);
    if (taken <= 0.0) return;
    c.co2_held[i] += taken;
    deposit_add(&C.deposit[idx], -taken / gv, C.scale);
    };
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) body(i);
}

// The CO2 flush (lifecycle.cu:235), death_kernel (:100-112) and mark_kernel
// (:119-135) in one pass: each touches only the cell's own state or, for the
// flush, its voxel through the exchange protocol of k_apply_awake.
template <class P>
__global__ __launch_bounds__(BLOCK) void k_life_mark(Cells<P> c, Counters* ctr, Grid C, int c_on,
                                                     unsigned long long* demand, const uint32_t* vox,
                                                     int disposition, double quota, int32_t* dflag, int32_t* drop) {
    const int n = ctr->nphys;
    for (int base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
    const int i = base + threadIdx.x;
    bool divides = false, dead = false;
    if (i < n) {
        if (c_on) {
            const uint32_t idx = vox[i];
            if (idx != NO_VOXEL) {
                const unsigned long long raw = atomicExch(&C.deposit[idx], 0ull);
                if (raw != 0ull) C.value[idx] += static_cast<float>(from_fixed(raw, C.scale));
                demand[idx] = 0ull;
            }
        }
        uint32_t f = c.flags[i];
        if ((f & ASTRO_FLAG_OCCUPIED) && (f & ASTRO_FLAG_ALIVE) &&
            static_cast<double>(c.t_local[i]) > ASTRO_CELL_LETHAL_TEMP) {
            f &= ~ASTRO_FLAG_ALIVE;
            c.flags[i] = f;
            c.death_cause[i] = ASTRO_DEATH_OVERHEATED;
            c.emit_power[i] = 0.0f;
            c.energy[i] = disposition == ASTRO_STORE_RETAIN ? c.energy[i] : 0.0;
        }
        const bool occ = (f & ASTRO_FLAG_OCCUPIED) != 0u, alive = (f & ASTRO_FLAG_ALIVE) != 0u;
        divides = occ && alive && c.co2_held[i] >= quota;
        dead = occ && !alive;
        const uint32_t s = c.lslot[i];
        if (s != NO_SLOT) {
            if (divides) dflag[s] = 1;          // cleared at the start of the tick
            if (!(occ && alive)) drop[s] = 1;   // compaction keeps occupied && alive
        }
    }
    wave_count_add(&ctr->births, divides);
    wave_count_add(&ctr->deads, dead);
    }
}

__global__ __launch_bounds__(BLOCK) void k_divide_scan(Counters* ctr, const int32_t* dflag, int32_t* prefix,
                                                       int32_t* total, ScanCtl sc) {
    const int n = ctr->births > 0 ? ctr->count : 0;
    if (n == 0) return;
    scan_exclusive(
        n, [&](int i) { return load4_i32(dflag, i, n); }, [&](int i, int4 ex) { store4_i32(prefix, i, n, ex); }, sc,
        total);
}

// divide_kernel (lifecycle.cu:140-204). The daughter's reference slot is
// count + prefix (the prefix taken over reference-slot order) and its id
// next_id + prefix; physically it is appended after the last entry.
template <class P>
__global__ __launch_bounds__(BLOCK) void k_divide(Cells<P> c, const Counters* ctr, const int32_t* prefix,
                                                  TickParams t) {
    if (ctr->births <= 0) return;
    const int np = ctr->nphys;
    const int n = ctr->count;
    const auto body = [&](const int i) {
    const uint32_t f = c.flags[i];
    if (!(f & ASTRO_FLAG_OCCUPIED) || !(f & ASTRO_FLAG_ALIVE)) return;
    if (!(c.co2_held[i] >= t.quota)) return;
    const int32_t pre = prefix[c.lslot[i]];
    const int32_t lslot = n + pre;
    if (lslot >= ctr->capacity) return;
    const int32_t slot = np + pre;
    c.lslot[slot] = static_cast<uint32_t>(lslot);
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
    const double s = sqrt(u * u < 1.0 ? 1.0 - u * u : 0.0);
    const double ax = s * cos(phi), ay = s * sin(phi), az = u;
    const double a = t.division_offset;
    const double px = static_cast<double>(c.x[i]), py = static_cast<double>(c.y[i]), pz = static_cast<double>(c.z[i]);
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
    c.rng_state[slot] = pcg_split(c.rng_state[i], did);
    c.taxis_memory[slot] = 0.0f;
    c.run_timer[slot] = 0.0f;
    c.t_local[slot] = c.t_local[i];
    c.irradiance[slot] = 0.0f;
    c.co2_local[slot] = c.co2_local[i];
    c.n2_local[slot] = c.n2_local[i];
    };
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < np; i += gridDim.x * BLOCK) body(i);
}

// The host bookkeeping of lifecycle_step (lifecycle.cu:251-294), on device.
__global__ void k_life_finalize(Counters* ctr, int compaction) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int32_t n = ctr->count, deads = ctr->deads;
    int32_t births = ctr->births;
    const int32_t new_deaths = deads - ctr->dead_slots_prev;
    if (new_deaths > 0) ctr->deaths_total += new_deaths;
    if (births > 0) {
        if (n + births > ctr->capacity) births = ctr->capacity - n;
        if (births > 0) {
            ctr->divisions_total += births;
            ctr->next_id += static_cast<uint64_t>(births);
            ctr->count = n + births;
            ctr->nphys += births;
        }
    }
    const bool compact = compaction && deads > 0;
    ctr->compacting = compact ? 1 : 0;
    ctr->dead_slots_prev = compact ? 0 : deads;
    ctr->births = 0;
    ctr->deads = 0;
}

// ---------------------------------------------------------------------------
// cell_store_compact (src/sim/cell_store.cu:414-456): survivors (occupied &&
// alive) keep their order and are renumbered 0..live-1 by an exclusive scan
// of the drop flags over reference-slot order; dropped entries become
// tombstones. Nothing is copied: storage order is the spatial sort order, and
// the next re-sort packs the tombstones out. All kernels return at once
// unless `compacting` is set.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_compact_scan(Counters* ctr, const int32_t* drop, int32_t* dprefix,
                                                        int32_t* total, ScanCtl sc) {
    const int n = ctr->compacting ? ctr->count : 0;
    if (n == 0) return;
    scan_exclusive(
        n, [&](int i) { return load4_i32(drop, i, n); }, [&](int i, int4 ex) { store4_i32(dprefix, i, n, ex); }, sc,
        total);
}

template <class P>
__global__ __launch_bounds__(BLOCK) void k_compact_renumber(Cells<P> c, const Counters* ctr, const int32_t* drop,
                                                            const int32_t* dprefix) {
    if (!ctr->compacting) return;
    const int np = ctr->nphys;
    const auto body = [&](const int i) {
    const uint32_t s = c.lslot[i];
    if (s == NO_SLOT) return;
    if (drop[s]) {
        c.lslot[i] = NO_SLOT;
        c.flags[i] = 0u;
    } else {
        c.lslot[i] = s - static_cast<uint32_t>(dprefix[s]);
    }
    };
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < np; i += gridDim.x * BLOCK) body(i);
}

__global__ void k_compact_finalize(Counters* ctr, const int32_t* total) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (ctr->compacting) {
        ctr->count -= *total;
        ctr->live = ctr->count;
        ctr->compacting = 0;
    }
}

// A direct compaction request (cell_store_compact outside lifecycle): mark
// the drop set afresh -- every live entry writes its own slot's flag -- and
// run the same three passes.
template <class P>
__global__ __launch_bounds__(BLOCK) void k_compact_mark(Cells<P> c, Counters* ctr, int32_t* drop) {
    const int np = ctr->nphys;
    const int gid = blockIdx.x * BLOCK + threadIdx.x;
    if (gid == 0) ctr->compacting = 1;
    for (int i = gid; i < np; i += gridDim.x * BLOCK) {
        const uint32_t s = c.lslot[i];
        if (s == NO_SLOT) continue;
        const uint32_t f = c.flags[i];
        drop[s] = ((f & ASTRO_FLAG_OCCUPIED) && (f & ASTRO_FLAG_ALIVE)) ? 0 : 1;
    }
}

// ---------------------------------------------------------------------------
// stats_cells_kernel (src/sim/stats.cu:27-43): per-element fixed point, wave
// reduction, one atomic per wave and word.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

template <class P>
__global__ __launch_bounds__(BLOCK) void k_stats_cells(Cells<P> c, const Counters* ctr, unsigned long long* acc) {
    const int n = ctr->nphys;
    for (int base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
    const int i = base + threadIdx.x;
    unsigned long long e = 0, tc = 0, na = 0, nd = 0, nw = 0, nm = 0;
    if (i < n) {
        const uint32_t f = c.flags[i];
        if (f & ASTRO_FLAG_OCCUPIED) {
            const bool alive = (f & ASTRO_FLAG_ALIVE) != 0u, awake = (f & ASTRO_FLAG_AWAKE) != 0u;
            na = alive; nd = !alive; nw = awake; nm = !awake;
            e = static_cast<unsigned long long>(to_fixed(c.energy[i], ASTRO_DEPOSIT_SCALE_STATS));
            if (alive) tc = static_cast<unsigned long long>(to_fixed(static_cast<double>(c.temp_cell[i]), ASTRO_DEPOSIT_SCALE_STATS));
        }
    }
    e = wave_sum_u64(e); tc = wave_sum_u64(tc); na = wave_sum_u64(na);
    nd = wave_sum_u64(nd); nw = wave_sum_u64(nw); nm = wave_sum_u64(nm);
    if ((threadIdx.x & 63) == 0) {
        if (e) atomicAdd(&acc[SW_ENERGY], e);
        if (tc) atomicAdd(&acc[SW_TEMP_CELL], tc);
        if (na) atomicAdd(&acc[SW_N_ALIVE], na);
        if (nd) atomicAdd(&acc[SW_N_DEAD], nd);
        if (nw) atomicAdd(&acc[SW_N_AWAKE], nw);
        if (nm) atomicAdd(&acc[SW_N_DORMANT], nm);
    }
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

template <class P>
void launch_thermal_substep(const Launch& L, const WorldDev<P>& w, double dt_sub, double cap, bool update_rec) {
    PROF_BEGIN("thermal");
    k_thermal<P><<<stride_grid(w.pcap), BLOCK, 0, L.stream>>>(w.c, w.ctr, w.g[ASTRO_FIELD_TEMPERATURE], dt_sub, cap,
                                                               w.rec, w.spos, update_rec ? 1 : 0);
    PROF_END();
}

template <class P>
void launch_apply_awake(const Launch& L, const WorldDev<P>& w, int field) {
    PROF_BEGIN("deposit_apply");
    k_apply_awake<P><<<stride_grid(w.pcap), BLOCK, 0, L.stream>>>(w.c, w.ctr, w.g[field]);
    PROF_END();
}

template <class P>
void launch_feed_taxis(const Launch& L, const WorldDev<P>& w, const TickParams& t, bool feed, bool taxis) {
    const int g = stride_grid(w.pcap);
    PROF_BEGIN(feed ? (taxis ? "feed_taxis" : "feed") : "taxis");
    if (feed && taxis)
        k_feed_taxis<P, true, true><<<g, BLOCK, 0, L.stream>>>(w.c, w.ctr, w.spos, w.dirs, t.ambient_irr, t.dt,
                                                               t.petrova_max_power);
    else if (feed)
        k_feed_taxis<P, true, false><<<g, BLOCK, 0, L.stream>>>(w.c, w.ctr, w.spos, w.dirs, t.ambient_irr, t.dt,
                                                                t.petrova_max_power);
    else if (taxis)
        k_feed_taxis<P, false, true><<<g, BLOCK, 0, L.stream>>>(w.c, w.ctr, w.spos, w.dirs, t.ambient_irr, t.dt,
                                                                t.petrova_max_power);
    PROF_END();
}

template <class P>
void launch_sample_base(const Launch& L, const WorldDev<P>& w, const TickParams& t) {
    PROF_BEGIN("sample_base");
    k_sample_base<P><<<stride_grid(w.pcap), BLOCK, 0, L.stream>>>(
        w.c, w.ctr, w.g[ASTRO_FIELD_TEMPERATURE], w.g[ASTRO_FIELD_CO2], w.g[ASTRO_FIELD_N2],
        w.on[ASTRO_FIELD_CO2] ? 1 : 0, w.on[ASTRO_FIELD_N2] ? 1 : 0, t.gravity_axis, w.spos, w.fs);
    PROF_END();
}

template <class P>
void launch_integrate(const Launch& L, const WorldDev<P>& w, const TickParams& t) {
    PROF_BEGIN("integrate");
    k_integrate<P><<<stride_grid(w.pcap), BLOCK, 0, L.stream>>>(w.c, w.ctr, w.spos, w.fs, t);
    PROF_END();
}

template <class P>
void launch_lifecycle(const Launch& L, const WorldDev<P>& w, const TickParams& t) {
    const int g = stride_grid(w.pcap);
    const bool c_on = w.on[ASTRO_FIELD_CO2];
    if (c_on) {
        PROF_BEGIN("co2_demand");
        k_demand<P><<<g, BLOCK, 0, L.stream>>>(w.c, w.ctr, w.g[ASTRO_FIELD_CO2], t.dt_bio, t.co2_gv, w.co2_demand,
                                                w.rin);
        PROF_END();
        PROF_BEGIN("co2_uptake");
        k_uptake<P><<<g, BLOCK, 0, L.stream>>>(w.c, w.ctr, w.g[ASTRO_FIELD_CO2], t.dt_bio, t.co2_gv, w.co2_demand,
                                                w.rin);
        PROF_END();
    }
    PROF_BEGIN("life_mark");
    k_life_mark<P><<<g, BLOCK, 0, L.stream>>>(w.c, w.ctr, w.g[ASTRO_FIELD_CO2], c_on ? 1 : 0, w.co2_demand, w.rin,
                                               t.store_disposition, t.quota, w.dflag, w.drop);
    PROF_END();
    PROF_BEGIN("divide_scan");
    k_divide_scan<<<scan_tiles_for(w.capacity), BLOCK, 0, L.stream>>>(w.ctr, w.dflag, w.prefix, w.scan_total, w.sc);
    PROF_END();
    PROF_BEGIN("divide");
    k_divide<P><<<g, BLOCK, 0, L.stream>>>(w.c, w.ctr, w.prefix, t);
    PROF_END();
    PROF_BEGIN("life_finalize");
    k_life_finalize<<<1, 64, 0, L.stream>>>(w.ctr, t.compaction);
    PROF_END();
    if (t.compaction) launch_compact<P>(L, w, false);
}

template <class P>
void launch_life_clear(const Launch& L, const WorldDev<P>& w) {
    hipMemsetAsync(w.dflag, 0, sizeof(int32_t) * w.capacity, L.stream);
    hipMemsetAsync(w.drop, 0, sizeof(int32_t) * w.capacity, L.stream);
}

template <class P>
void launch_compact(const Launch& L, const WorldDev<P>& w, bool forced) {
    const int g = stride_grid(w.pcap);
    if (forced) {
        PROF_BEGIN("compact_mark");
        k_compact_mark<P><<<g, BLOCK, 0, L.stream>>>(w.c, w.ctr, w.drop);
        PROF_END();
    }
    PROF_BEGIN("compact_scan");
    k_compact_scan<<<scan_tiles_for(w.capacity), BLOCK, 0, L.stream>>>(w.ctr, w.drop, w.dprefix, w.scan_total, w.sc);
    PROF_END();
    PROF_BEGIN("compact_renumber");
    k_compact_renumber<P><<<g, BLOCK, 0, L.stream>>>(w.c, w.ctr, w.drop, w.dprefix);
    PROF_END();
    PROF_BEGIN("compact_finalize");
    k_compact_finalize<<<1, 64, 0, L.stream>>>(w.ctr, w.scan_total);
    PROF_END();
}

template <class P>
void launch_stats_cells(const Launch& L, const WorldDev<P>& w) {
    k_stats_cells<P><<<stride_grid(w.pcap), BLOCK, 0, L.stream>>>(w.c, w.ctr, w.stats);
}

#define INSTANTIATE(P)                                                                                          \
    template void launch_spawn<P>(const Launch&, const WorldDev<P>&, const SpawnArgs&, int32_t, int32_t,       \
                                  uint64_t, uint64_t);                                                          \
    template void launch_thermal_substep<P>(const Launch&, const WorldDev<P>&, double, double, bool);          \
    template void launch_apply_awake<P>(const Launch&, const WorldDev<P>&, int);                                \
    template void launch_feed_taxis<P>(const Launch&, const WorldDev<P>&, const TickParams&, bool, bool);       \
    template void launch_sample_base<P>(const Launch&, const WorldDev<P>&, const TickParams&);                  \
    template void launch_integrate<P>(const Launch&, const WorldDev<P>&, const TickParams&);                    \
    template void launch_lifecycle<P>(const Launch&, const WorldDev<P>&, const TickParams&);                    \
    template void launch_compact<P>(const Launch&, const WorldDev<P>&, bool);                                    \
    template void launch_life_clear<P>(const Launch&, const WorldDev<P>&);                                      \
    template void launch_stats_cells<P>(const Launch&, const WorldDev<P>&);

INSTANTIATE(double)
INSTANTIATE(float)

}  // namespace astro_dev
