// world.hip -- the C ABI of include/astro_hot.h: world lifetime, host<->device
// copies, the tick sequence of src/sim/step.cu:109-184, graph capture and
// per-kernel timing.
//
// A tick never synchronizes: the population count, next id and per-tick
// tallies live in the device Counters block, every grid is sized for the
// capacity, and each kernel reads the live count itself. astro_world_step
// therefore captures one tick (two, when a field's ping-pong parity would not
// return after one) into a hipGraph and replays it.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <span>
#include <string>
#include <type_traits>
#include <vector>

#include "astro_hot.h"
#include "kernels.h"

using namespace astro_dev;

namespace {

// ---------------------------------------------------------------------------
// Field bookkeeping (src/fields/diffuse.cu:91-119), extended to nz planes.
// ---------------------------------------------------------------------------
struct FieldHost {
    bool on = false, square2d = false;
    int nx = 0, ny = 0, nz = 0;
    double dx = 0, dy = 0, dz = 0, D = 0, ambient = 0, scale = 1;
    int bc = ASTRO_BC_NEUMANN;
    double rob_x = 0, rob_y = 0, rob_z = 0;
    float* value = nullptr;
    float* scratch = nullptr;
    unsigned long long* deposit = nullptr;
    size_t count() const { return on ? static_cast<size_t>(nx) * ny * nz : 0; }
    Grid grid() const {
        Grid g{};
        g.value = value;
        g.scratch = scratch;
        g.deposit = deposit;
        g.nx = nx; g.ny = ny; g.nz = nz;
        g.fdx = static_cast<float>(dx);
        g.fdy = static_cast<float>(dy);
        g.fdz = static_cast<float>(dz);
        g.scale = scale;
        return g;
    }
};

// The ping-pong buffers of every field, as a tick leaves them on the host.
struct BufPtrs {
    float* value[ASTRO_FIELD_COUNT];
    float* scratch[ASTRO_FIELD_COUNT];
};

void field_init(FieldHost& g, astro_grid_dims d, double w, double h, double depth, double D, double ambient,
                double scale, int bc) {
    g.on = d.nx > 0;
    if (!g.on) return;
    g.nx = d.nx; g.ny = d.ny; g.nz = d.nz;
    g.square2d = (g.nz == 1 && g.nx == g.ny);
    g.dx = w / static_cast<double>(g.nx);
    g.dy = g.square2d ? g.dx : h / static_cast<double>(g.ny);
    g.dz = depth / static_cast<double>(g.nz);
    g.D = D; g.ambient = ambient; g.scale = scale; g.bc = bc;
    g.rob_x = g.dx * ASTRO_AIR_CONVECTION_H / ASTRO_WATER_CONDUCTIVITY;
    g.rob_y = g.dy * ASTRO_AIR_CONVECTION_H / ASTRO_WATER_CONDUCTIVITY;
    g.rob_z = g.dz * ASTRO_AIR_CONVECTION_H / ASTRO_WATER_CONDUCTIVITY;
}

// substeps_for_dt (src/fields/grid.cuh:51-53); 3D uses the 7-point bound.
int substeps_for_dt(const FieldHost& g, double dt) {
    if (g.square2d) return static_cast<int>(dt / (g.dx * g.dx / (4.0 * g.D))) + 1;
    double inv = 1.0 / (g.dx * g.dx) + 1.0 / (g.dy * g.dy);
    if (g.nz > 1) inv += 1.0 / (g.dz * g.dz);
    return static_cast<int>(dt / (1.0 / (2.0 * g.D * inv))) + 1;
}

// Coefficients of one substep; false when unstable (diffuse.cu:146, :172).
bool ftcs_coeffs(const FieldHost& g, double dt_sub, Ftcs& f) {
    f = Ftcs{};
    f.square2d = g.square2d;
    f.bc = g.bc;
    f.amb = static_cast<float>(g.ambient);
    f.rx = static_cast<float>(g.rob_x);
    f.ry = static_cast<float>(g.rob_y);
    f.rz = static_cast<float>(g.rob_z);
    if (g.square2d) {
        f.coeff = static_cast<float>(g.D * dt_sub / (g.dx * g.dx));
        return f.coeff <= 0.25f;
    }
    f.cx = static_cast<float>(g.D * dt_sub / (g.dx * g.dx));
    f.cy = static_cast<float>(g.D * dt_sub / (g.dy * g.dy));
    f.cz = g.nz > 1 ? static_cast<float>(g.D * dt_sub / (g.dz * g.dz)) : 0.0f;
    return static_cast<double>(f.cx) + f.cy + f.cz <= 0.5;
}

// HIP-event timing of every launch on the world's stream.
struct EventProfiler final : Profiler {
    hipStream_t stream = nullptr;
    struct Pending { std::string name; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<int64_t, double>> totals;
    std::vector<std::string> order;
    Pending cur;

    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        hipEventCreate(&e);
        return e;
    }
    void begin(const char* name) override {
        cur.name = name;
        cur.a = take();
        cur.b = take();
        hipEventRecord(cur.a, stream);
    }
    void end() override {
        hipEventRecord(cur.b, stream);
        pending.push_back(cur);
    }
    void drain() {
        hipStreamSynchronize(stream);
        for (Pending& p : pending) {
            float ms = 0.0f;
            hipEventElapsedTime(&ms, p.a, p.b);
            auto it = totals.find(p.name);
            if (it == totals.end()) {
                order.push_back(p.name);
                it = totals.emplace(p.name, std::make_pair<int64_t, double>(0, 0.0)).first;
            }
            it->second.first += 1;
            it->second.second += ms;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    ~EventProfiler() override {
        for (Pending& p : pending) { pool.push_back(p.a); pool.push_back(p.b); }
        for (hipEvent_t e : pool) hipEventDestroy(e);
    }
};

// One device allocation carved into 256 B-aligned arrays
// (the reference's carve-once blob, src/sim/cell_store.cu:139-248).
struct Carver {
    char* base = nullptr;
    size_t off = 0;
    template <class T>
    T* take(size_t n) {
        const size_t bytes = ((sizeof(T) * n + 255) / 256) * 256;
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += bytes ? bytes : 256;
        return p;
    }
};

template <class P>
void carve_cells(Carver& cv, Cells<P>& c, size_t cap) {
    c.lslot = cv.take<uint32_t>(cap);
    c.id = cv.take<uint64_t>(cap); c.flags = cv.take<uint32_t>(cap); c.death_cause = cv.take<uint8_t>(cap);
    c.x = cv.take<P>(cap); c.y = cv.take<P>(cap); c.z = cv.take<P>(cap);
    c.vx = cv.take<P>(cap); c.vy = cv.take<P>(cap); c.vz = cv.take<P>(cap);
    c.energy = cv.take<double>(cap); c.temp_cell = cv.take<float>(cap); c.emit_power = cv.take<float>(cap);
    c.dir_x = cv.take<float>(cap); c.dir_y = cv.take<float>(cap); c.dir_z = cv.take<float>(cap);
    c.biomass = cv.take<double>(cap); c.co2_held = cv.take<double>(cap); c.age_s = cv.take<float>(cap);
    c.rng_state = cv.take<uint64_t>(cap); c.taxis_memory = cv.take<float>(cap); c.run_timer = cv.take<float>(cap);
    c.t_local = cv.take<float>(cap); c.irradiance = cv.take<float>(cap); c.co2_local = cv.take<float>(cap);
    c.n2_local = cv.take<float>(cap);
}

}  // namespace

struct astro_world {
    astro_world_desc desc{};
    hipStream_t stream = nullptr;
    int device = 0;
    int32_t capacity = 0;
    bool fp32 = false;
    void* pool = nullptr;
    size_t pool_bytes = 0;

    Cells<double> c64[2]{};
    Cells<float> c32[2]{};
    Counters* ctr = nullptr;
    FieldHost f[ASTRO_FIELD_COUNT];

    HashGeo hg{};
    double hcell = 0.0;
    int32_t pcap = 0;
    int32_t resort_period = 16;
    uint32_t *keys = nullptr, *rin = nullptr, *bcount = nullptr, *vals_sorted = nullptr;
    uint32_t *tphys = nullptr, *sphys = nullptr;
    int32_t *bstart = nullptr, *spos = nullptr, *dflag = nullptr, *prefix = nullptr;
    int32_t *drop = nullptr, *dprefix = nullptr, *bsum = nullptr, *dsum = nullptr, *xsum = nullptr;
    uint32_t* dlist = nullptr;
    float* co2_avail = nullptr;
    void *trec = nullptr, *rec = nullptr;
    void* fs = nullptr;              // Force<double> or Force<float> per layout
    double* dirs = nullptr;
    uint32_t* co2_demand = nullptr;
    uint32_t* tcount = nullptr;   // live cells per temperature voxel, during the thermal step
    uint8_t* talone = nullptr;    // per physical entry: alone in its voxel this tick
    // The thermostat's deposit path (astro_world_set_deposit_path) and the
    // host's estimate of the live and awake populations that the automatic
    // choice weighs, refreshed by spawn, upload and stats.
    int32_t deposit_path = ASTRO_DEPOSIT_AUTO;
    int64_t live_hint = 0, awake_hint = 0;
    unsigned long long* stats = nullptr;

    double physics_rate = 1.0, biology_rate = 1.0;
    uint64_t tick = 0;
    double sim_time = 0.0;
    int64_t div_reported = 0, deaths_reported = 0;
    bool hash_valid = false;
    // The sticky non_canon_run (src/sim/world.cuh:157-158), restored from
    // a snapshot header; parameter overrides also make a run non-canon.
    bool non_canon = false;

    // World::flash_active (src/sim/world.cuh:178-179) and its audit words.
    bool flash_on = false;
    unsigned long long* flash_accum = nullptr;

    // The Taumoeba store; the tick runs predation once any were added.
    TauDev tau{};
    int32_t tau_cap = 0;
    bool tau_on = false;

    // Captured ticks. A field with an odd number of substeps per tick ends
    // each tick with its ping-pong buffers swapped, and a captured launch
    // bakes the buffers it read: so the graphs are kept per phase of that
    // swap. graph[p][t - 1] runs t ticks (t = 1, 2) starting in phase p;
    // phase 0 is the arrangement at capture, and a tick flips the phase when
    // some field's substep count is odd. Without such a field only
    // graph[0][0] exists and the phase stays 0.
    hipGraph_t graph[2][2] = {};
    hipGraphExec_t gexec[2][2] = {};
    BufPtrs gstart[2] = {};   // the buffers each phase starts from
    bool graph_ok = false;
    int phase = 0;
    // Bumped by every change that drops the captured tick (clock, light,
    // flash, stage calls, field swaps, ...). An ensemble that captured this
    // world's launch arguments compares it to know when to capture again.
    uint64_t generation = 0;

    std::unique_ptr<EventProfiler> prof;
    bool prof_on = false;
    std::string err;
};

namespace {

int32_t fail(astro_world* w, int32_t st, const std::string& what) {
    if (w) w->err = what;
    return st;
}

int32_t hip_check(astro_world* w, hipError_t e, const char* what) {
    if (e == hipSuccess) return ASTRO_OK;
    return fail(w, e == hipErrorOutOfMemory ? ASTRO_OUT_OF_MEMORY : ASTRO_DEVICE_ERROR,
                std::string(what) + ": " + hipGetErrorString(e));
}

template <class P>
WorldDev<P> view(astro_world* w);

template <>
WorldDev<double> view<double>(astro_world* w) {
    WorldDev<double> d{};
    d.c = w->c64[0];
    d.alt = w->c64[1];
    d.rec = static_cast<Rec<double>*>(w->rec);
    d.trec = static_cast<Rec<double>*>(w->trec);
    return d;
}
template <>
WorldDev<float> view<float>(astro_world* w) {
    WorldDev<float> d{};
    d.c = w->c32[0];
    d.alt = w->c32[1];
    d.rec = static_cast<Rec<float>*>(w->rec);
    d.trec = static_cast<Rec<float>*>(w->trec);
    return d;
}

template <class P>
WorldDev<P> make_view(astro_world* w) {
    WorldDev<P> d = view<P>(w);
    d.ctr = w->ctr;
    d.capacity = w->capacity;
    d.pcap = w->pcap;
    d.resort_period = w->resort_period;
    for (int i = 0; i < ASTRO_FIELD_COUNT; ++i) {
        d.g[i] = w->f[i].grid();
        d.on[i] = w->f[i].on;
    }
    d.hg = w->hg;
    d.keys = w->keys; d.rin = w->rin; d.bcount = w->bcount; d.bstart = w->bstart;
    d.vals_sorted = w->vals_sorted; d.spos = w->spos; d.fs = static_cast<Force<P>*>(w->fs); d.dirs = w->dirs;
    d.tphys = w->tphys; d.sphys = w->sphys;
    d.dflag = w->dflag; d.prefix = w->prefix; d.co2_demand = w->co2_demand; d.tcount = w->tcount; d.talone = w->talone;
    d.drop = w->drop; d.dprefix = w->dprefix; d.bsum = w->bsum; d.dsum = w->dsum; d.xsum = w->xsum;
    d.dlist = w->dlist;
    d.co2_avail = w->co2_avail;
    d.stats = w->stats;
    d.flash = w->flash_accum;
    return d;
}

Launch launch_of(astro_world* w) { return Launch{w->stream, w->prof_on ? w->prof.get() : nullptr, nullptr}; }

TickParams tick_params(const astro_world* w, double dt) {
    const astro_world_desc& d = w->desc;
    TickParams t{};
    t.dt = dt;
    t.dt_bio = dt * w->biology_rate;
    t.chamber_w = d.chamber_w; t.chamber_h = d.chamber_h; t.chamber_d = d.chamber_d;
    t.boundary_x = d.boundary_x; t.boundary_y = d.boundary_y; t.gravity_axis = d.gravity_axis;
    t.thermal_noise = d.thermal_noise; t.contact = d.contact_enabled; t.adhesion = d.adhesion_enabled;
    t.occlusion_exact = d.occlusion_exact;
    t.store_disposition = d.store_disposition; t.compaction = d.compaction_enabled;
    t.k_contact = ASTRO_CONTACT_STIFFNESS / w->physics_rate;
    t.petrova_max_power = d.petrova_max_power;
    t.petrova_flash_power = d.petrova_flash_power;
    t.quota = d.co2_mass_per_division;
    t.division_offset = d.division_offset;
    t.source = d.light_enabled ? d.light_irradiance : 0.0f;
    t.ambient_irr = d.ambient_irradiance;
    // normalize(Vec3{dir_x, dir_y, 0}) (src/sim/emission.cu:158)
    const double lx = d.light_dir_x, ly = d.light_dir_y, lz = 0.0;
    const double l = std::sqrt(lx * lx + ly * ly + lz * lz);
    if (l > 0.0) {
        const double inv = 1.0 / l;
        t.ldx = lx * inv; t.ldy = ly * inv; t.ldz = lz * inv;
    } else {
        t.ldx = 0.0; t.ldy = 0.0; t.ldz = 1.0;
    }
    t.sweep_axis = (std::fabs(t.ldx) >= std::fabs(t.ldy)) ? 0 : 1;
    t.sweep_sign = t.sweep_axis == 0 ? (t.ldx >= 0.0 ? 1 : -1) : (t.ldy >= 0.0 ? 1 : -1);
    const FieldHost& I = w->f[ASTRO_FIELD_IRRADIANCE];
    t.face_area = (t.sweep_axis == 0 ? I.dy : I.dx) * I.dz;
    const FieldHost& T = w->f[ASTRO_FIELD_TEMPERATURE];
    t.T_cap = ASTRO_WATER_DENSITY * ASTRO_WATER_SPECIFIC_HEAT * T.dx * T.dy * T.dz;
    const FieldHost& C = w->f[ASTRO_FIELD_CO2];
    t.co2_gv = C.dx * C.dy * C.dz;
    t.rng_seed = d.seed;
    t.philox = d.rng_kind == ASTRO_RNG_PHILOX ? 1 : 0;
    return t;
}

bool lit(const astro_world* w) {
    const float source = w->desc.light_enabled ? w->desc.light_irradiance : 0.0f;
    return !(source <= 0.0f && w->desc.ambient_irradiance <= 0.0f);
}

// The worlds one enqueue covers: a single world, or the chambers of an
// ensemble stepped in lock step. Host-side decisions (which stages run, substep
// counts) are taken from the first world; an ensemble checks before each step
// that every chamber agrees (tick_shape).
using WorldSpan = std::span<astro_world* const>;

template <class P>
std::vector<WorldDev<P>> views_of(WorldSpan ws) {
    std::vector<WorldDev<P>> v;
    v.reserve(ws.size());
    for (astro_world* w : ws) v.push_back(make_view<P>(w));
    return v;
}

std::vector<TickParams> ticks_of(WorldSpan ws, double dt) {
    std::vector<TickParams> t;
    t.reserve(ws.size());
    for (astro_world* w : ws) t.push_back(tick_params(w, dt));
    return t;
}

TauTick tau_tick(const astro_world* w, double dt) {
    const astro_world_desc& d = w->desc;
    TauTick t{};
    t.dt = dt;
    t.dt_bio = dt * w->biology_rate;
    t.ambient_temp = d.ambient_temp;
    t.chamber_w = d.chamber_w;
    t.chamber_h = d.chamber_h;
    t.chamber_d = d.chamber_d;
    t.boundary_x = d.boundary_x;
    t.boundary_y = d.boundary_y;
    t.thermal_noise = d.thermal_noise;
    t.store_disposition = d.store_disposition;
    t.compaction = d.tau_compaction_enabled;
    return t;
}

// grid_diffuse / grid_diffuse_substep: enqueue substeps, swapping on the host.
int32_t enqueue_diffuse(WorldSpan ws, const Launch& L, int field, double dt) {
    if (!ws[0]->f[field].on) return ASTRO_OK;
    const int steps = substeps_for_dt(ws[0]->f[field], dt);
    if (steps <= 0) return ASTRO_OK;
    const double dt_sub = dt / steps;
    std::vector<Ftcs> fc(ws.size());
    std::vector<Grid> gs(ws.size());
    for (size_t i = 0; i < ws.size(); ++i)
        if (!ftcs_coeffs(ws[i]->f[field], dt_sub, fc[i]))
            return fail(ws[i], ASTRO_INVALID_ARGUMENT, "diffusion coefficient unstable");
    static const char* names[ASTRO_FIELD_COUNT] = {"ftcs_T", "ftcs_CO2", "ftcs_N2", "ftcs_irradiance"};
    for (int s = 0; s < steps; ++s) {
        for (size_t i = 0; i < ws.size(); ++i) gs[i] = ws[i]->f[field].grid();
        launch_ftcs(L, gs, fc, names[field]);
        for (astro_world* w : ws) std::swap(w->f[field].value, w->f[field].scratch);
    }
    return ASTRO_OK;
}

// Which thermostat deposit path a tick takes (the field is the same bit for
// bit either way). Counting the occupied entries per voxel costs about three
// memory operations per entry per tick; each awake cell alone in its voxel
// then saves an atomic add and an exchange per substep.
bool thermal_direct(const astro_world* w, int substeps) {
    if (w->deposit_path == ASTRO_DEPOSIT_ATOMIC) return false;
    if (w->deposit_path == ASTRO_DEPOSIT_DIRECT) return true;
    return 2 * w->awake_hint * substeps > 3 * w->live_hint;
}

// thermal_step (src/sim/thermal.cu:75-103).
template <class P>
// counted: the tick's key pass has already counted the entries per voxel.
// stamp: the last substep also stamps the occluders for the emission stage.
int32_t enqueue_thermal(WorldSpan ws, const Launch& L, double dt, Ticks ts, bool counted = false,
                        bool stamp = false) {
    const int substeps = substeps_for_dt(ws[0]->f[ASTRO_FIELD_TEMPERATURE], dt);
    const double dt_sub = dt / substeps;
    std::vector<Ftcs> fc(ws.size());
    std::vector<Grid> gs(ws.size());
    for (size_t i = 0; i < ws.size(); ++i)
        if (!ftcs_coeffs(ws[i]->f[ASTRO_FIELD_TEMPERATURE], dt_sub, fc[i]))
            return fail(ws[i], ASTRO_INVALID_ARGUMENT, "diffusion coefficient unstable");
    const bool direct = thermal_direct(ws[0], substeps);
    if (direct && !counted) launch_thermal_count<P>(L, views_of<P>(ws));
    for (int s = 0; s < substeps; ++s) {
        const std::vector<WorldDev<P>> d = views_of<P>(ws);
        const int mode = !direct ? 0 : (s == 0 ? 1 : 2);
        launch_thermal_substep<P>(L, d, ts, dt_sub, ws[0]->hash_valid, mode, s, s == substeps - 1, stamp);
        launch_apply_awake<P>(L, d, ASTRO_FIELD_TEMPERATURE, mode, s);
        for (size_t i = 0; i < ws.size(); ++i) gs[i] = ws[i]->f[ASTRO_FIELD_TEMPERATURE].grid();
        launch_ftcs(L, gs, fc, "ftcs_T");
        for (astro_world* w : ws) std::swap(w->f[ASTRO_FIELD_TEMPERATURE].value, w->f[ASTRO_FIELD_TEMPERATURE].scratch);
    }
    return ASTRO_OK;
}

// emission_step, directional path (src/sim/emission.cu:145-176), minus the
// absorption which rides in the feed/taxis kernel. Only chambers with a
// positive source stamp occluders.
// stamped: the thermal step's last substep has stamped the occluders already.
template <class P>
void enqueue_light(WorldSpan ws, const Launch& L, Ticks ts, bool stamped_already = false) {
    const std::vector<WorldDev<P>> d = views_of<P>(ws);
    std::vector<WorldDev<P>> stamped;
    std::vector<Grid> irr;
    for (size_t i = 0; i < ws.size(); ++i) {
        if (ts[i].source > 0.0f) stamped.push_back(d[i]);
        irr.push_back(ws[i]->f[ASTRO_FIELD_IRRADIANCE].grid());
    }
    if (!stamped.empty() && !stamped_already) launch_stamp<P>(L, stamped);
    launch_sweep(L, irr, ts);
    // crowded: at least one live cell per hash bucket on average, in a launch
    // whose sorted records all fit in one XCD's 4 MiB L2 (the row walk reads
    // its neighbours from global memory, one dependent load after another)
    int64_t rec_bytes = 0;
    for (astro_world* w : ws) rec_bytes += static_cast<int64_t>(w->capacity) * static_cast<int64_t>(sizeof(Rec<P>));
    std::vector<char> crowded;
    for (astro_world* w : ws) crowded.push_back(w->live_hint >= w->hg.nb && rec_bytes <= (int64_t{4} << 20));
    launch_feed_walk<P>(L, d, ts, crowded);
}

// One tick of world_step (src/sim/step.cu:109-184). Host clock untouched.
template <class P>
int32_t enqueue_tick(WorldSpan ws, const Launch& L) {
    const double dt = ASTRO_DT_PHYSICS * ws[0]->physics_rate;
    const std::vector<TickParams> ts = ticks_of(ws, dt);
    const astro_world_desc& ds = ws[0]->desc;
    // When the thermal step takes the direct deposit path, its per-voxel
    // counts ride in the key pass, which reads the same positions.
    const bool count_t =
        ds.thermal_enabled && thermal_direct(ws[0], substeps_for_dt(ws[0]->f[ASTRO_FIELD_TEMPERATURE], dt));
    launch_hash_build<P>(L, views_of<P>(ws), true, count_t);
    launch_resort<P>(L, views_of<P>(ws));
    for (astro_world* w : ws) w->hash_valid = true;
    int32_t st;
    // A lit tick's occluder stamp rides in the thermal step's last substep,
    // which reads every entry's flags as the emission stage needs them.
    const bool feed = ds.emission_enabled && lit(ws[0]);
    const bool stamp_t = feed && ds.thermal_enabled;
    if (ds.thermal_enabled && (st = enqueue_thermal<P>(ws, L, dt, ts, count_t, stamp_t)) != ASTRO_OK) return st;
    if (feed) enqueue_light<P>(ws, L, ts, stamp_t);
    if (ws[0]->flash_on) {
        // Stage 9b (step.cu:158-162): the armed flash runs instead of taxis,
        // after the feed and before the motion reads emit_power.
        if (feed) launch_feed_taxis<P>(L, views_of<P>(ws), ts, true, false);
        launch_flash<P>(L, views_of<P>(ws), ts);
        launch_cell_forces<P>(L, views_of<P>(ws), ts, false, false);
    } else {
        launch_cell_forces<P>(L, views_of<P>(ws), ts, feed, ds.taxis_enabled != 0);
    }
    // Stages 7-8 (step.cu:168-171). Both accumulators are empty here: CO2
    // uptake is flushed inside lifecycle and nothing in this path deposits N2,
    // so the reference's dense flush is a no-op and is not issued. The CO2
    // substeps run before the integrator instead of after it: nothing between
    // reads the CO2 field (the sample above has already taken its values), and
    // the integrate pass can then book the lifecycle's CO2 demand itself.
    if ((st = enqueue_diffuse(ws, L, ASTRO_FIELD_CO2, dt)) != ASTRO_OK) return st;
    // With predators the demand is booked inside the lifecycle, after the
    // engulfs: a cell eaten this tick asks for no CO2.
    const bool predators = ws[0]->tau_on;
    const bool book = ds.lifecycle_enabled && ws[0]->f[ASTRO_FIELD_CO2].on && !predators;
    launch_integrate<P>(L, views_of<P>(ws), ts, book);
    if ((st = enqueue_diffuse(ws, L, ASTRO_FIELD_N2, dt)) != ASTRO_OK) return st;
    // Stage 10 (step.cu:172-175), on the hash of stage 1.
    if (predators)
        for (astro_world* w : ws) launch_predation<P>(L, make_view<P>(w), w->tau, tau_tick(w, dt));
    if (ds.lifecycle_enabled) launch_lifecycle<P>(L, views_of<P>(ws), ts, book);
    for (astro_world* w : ws) w->hash_valid = false;
    return ASTRO_OK;
}

int32_t tick_once(WorldSpan ws, const Launch& L) {
    return ws[0]->fp32 ? enqueue_tick<float>(ws, L) : enqueue_tick<double>(ws, L);
}

int32_t tick_once(astro_world* w, const Launch& L) {
    astro_world* const one_w[1] = {w};
    return tick_once(WorldSpan(one_w, 1), L);
}

void advance_clock(astro_world* w, int32_t ticks) {
    const double dt = ASTRO_DT_PHYSICS * w->physics_rate;
    for (int32_t i = 0; i < ticks; ++i) {
        ++w->tick;
        w->sim_time += dt;
    }
}

// Ticks per captured graph: two when some field's ping-pong ends swapped.
int ticks_per_graph(const astro_world* w) {
    const double dt = ASTRO_DT_PHYSICS * w->physics_rate;
    int swaps[3] = {0, 0, 0};
    if (w->desc.thermal_enabled) swaps[0] = substeps_for_dt(w->f[ASTRO_FIELD_TEMPERATURE], dt);
    if (w->f[ASTRO_FIELD_CO2].on) swaps[1] = substeps_for_dt(w->f[ASTRO_FIELD_CO2], dt);
    if (w->f[ASTRO_FIELD_N2].on) swaps[2] = substeps_for_dt(w->f[ASTRO_FIELD_N2], dt);
    for (int s : swaps)
        if (s & 1) return 2;
    return 1;
}

void drop_graph(astro_world* w) {
    ++w->generation;
    for (int p = 0; p < 2; ++p)
        for (int t = 0; t < 2; ++t) {
            if (w->gexec[p][t]) hipGraphExecDestroy(w->gexec[p][t]);
            if (w->graph[p][t]) hipGraphDestroy(w->graph[p][t]);
            w->gexec[p][t] = nullptr;
            w->graph[p][t] = nullptr;
        }
    w->graph_ok = false;
    w->phase = 0;
}

BufPtrs bufs_of(const astro_world* w) {
    BufPtrs b{};
    for (int f = 0; f < ASTRO_FIELD_COUNT; ++f) {
        b.value[f] = w->f[f].value;
        b.scratch[f] = w->f[f].scratch;
    }
    return b;
}

void set_bufs(astro_world* w, const BufPtrs& b) {
    for (int f = 0; f < ASTRO_FIELD_COUNT; ++f) {
        w->f[f].value = b.value[f];
        w->f[f].scratch = b.scratch[f];
    }
}

int32_t check_device_errors(astro_world* w) {
    Counters c{};
    int32_t st = hip_check(w, hipMemcpy(&c, w->ctr, sizeof(c), hipMemcpyDeviceToHost), "read counters");
    if (st != ASTRO_OK) return st;
    if (c.error_flags & ERR_SCAN) return fail(w, ASTRO_DEVICE_ERROR, "prefix scan tile total mismatch");
    return ASTRO_OK;
}

int32_t read_counters(astro_world* w, Counters& c) {
    int32_t st = hip_check(w, hipStreamSynchronize(w->stream), "synchronize");
    if (st != ASTRO_OK) return st;
    st = hip_check(w, hipMemcpy(&c, w->ctr, sizeof(c), hipMemcpyDeviceToHost), "read counters");
    if (st != ASTRO_OK) return st;
    if (c.error_flags & ERR_SCAN) return fail(w, ASTRO_DEVICE_ERROR, "prefix scan tile total mismatch");
    return ASTRO_OK;
}

int32_t write_counters(astro_world* w, const Counters& c) {
    int32_t st = hip_check(w, hipMemcpyAsync(w->ctr, &c, sizeof(c), hipMemcpyHostToDevice, w->stream), "write counters");
    if (st != ASTRO_OK) return st;
    return hip_check(w, hipStreamSynchronize(w->stream), "synchronize");
}

// Host <-> device copies of one array, converting the position width.
template <class D, class H>
int32_t put(astro_world* w, D* dst, const H* src, int32_t n) {
    if (!src || n <= 0) return ASTRO_OK;
    if constexpr (std::is_same<D, H>::value) {
        return hip_check(w, hipMemcpy(dst, src, sizeof(D) * n, hipMemcpyHostToDevice), "upload");
    } else {
        std::vector<D> tmp(n);
        for (int32_t i = 0; i < n; ++i) tmp[i] = static_cast<D>(src[i]);
        return hip_check(w, hipMemcpy(dst, tmp.data(), sizeof(D) * n, hipMemcpyHostToDevice), "upload");
    }
}
template <class D, class H>
int32_t get(astro_world* w, H* dst, const D* src, int32_t n) {
    if (!dst || n <= 0) return ASTRO_OK;
    if constexpr (std::is_same<D, H>::value) {
        return hip_check(w, hipMemcpy(dst, src, sizeof(D) * n, hipMemcpyDeviceToHost), "download");
    } else {
        std::vector<D> tmp(n);
        int32_t st = hip_check(w, hipMemcpy(tmp.data(), src, sizeof(D) * n, hipMemcpyDeviceToHost), "download");
        if (st != ASTRO_OK) return st;
        for (int32_t i = 0; i < n; ++i) dst[i] = static_cast<H>(tmp[i]);
        return ASTRO_OK;
    }
}

template <class P>
int32_t upload_cells(astro_world* w, const Cells<P>& c, const astro_cell_soa* s, int32_t n) {
    int32_t st = ASTRO_OK;
    // Uploaded cells are stored in reference order: physical == logical.
    std::vector<uint32_t> iota(n);
    for (int32_t i = 0; i < n; ++i) iota[i] = static_cast<uint32_t>(i);
    if ((st = put(w, c.lslot, iota.data(), n)) != ASTRO_OK) return st;
#define PUT(a) \
    if ((st = put(w, c.a, s->a, n)) != ASTRO_OK) return st;
    PUT(id) PUT(flags) PUT(death_cause) PUT(x) PUT(y) PUT(z) PUT(vx) PUT(vy) PUT(vz) PUT(energy) PUT(temp_cell)
    PUT(emit_power) PUT(dir_x) PUT(dir_y) PUT(dir_z) PUT(biomass) PUT(co2_held) PUT(age_s) PUT(rng_state)
    PUT(taxis_memory) PUT(run_timer) PUT(t_local) PUT(irradiance) PUT(co2_local) PUT(n2_local)
#undef PUT
    return st;
}

// Device array over physical entries [0, np) -> host array in reference-slot
// order, entries [0, m).
template <class D, class H>
int32_t get_logical(astro_world* w, H* dst, const D* src, int32_t np, const std::vector<uint32_t>& lslot, int32_t m) {
    if (!dst || np <= 0) return ASTRO_OK;
    std::vector<D> tmp(np);
    int32_t st = hip_check(w, hipMemcpy(tmp.data(), src, sizeof(D) * np, hipMemcpyDeviceToHost), "download");
    if (st != ASTRO_OK) return st;
    for (int32_t p = 0; p < np; ++p) {
        const uint32_t s = lslot[p];
        if (s != NO_SLOT && s < static_cast<uint32_t>(m)) dst[s] = static_cast<H>(tmp[p]);
    }
    return ASTRO_OK;
}

template <class P>
int32_t download_cells(astro_world* w, const Cells<P>& c, astro_cell_soa* s, int32_t n) {
    int32_t st = ASTRO_OK;
    Counters k{};
    if ((st = hip_check(w, hipMemcpy(&k, w->ctr, sizeof(k), hipMemcpyDeviceToHost), "counters")) != ASTRO_OK)
        return st;
    const int32_t np = k.nphys;
    std::vector<uint32_t> lslot(np > 0 ? np : 0);
    if (np > 0 &&
        (st = hip_check(w, hipMemcpy(lslot.data(), c.lslot, sizeof(uint32_t) * np, hipMemcpyDeviceToHost), "lslot")) != ASTRO_OK)
        return st;
#define GET(a) \
    if ((st = get_logical(w, s->a, c.a, np, lslot, n)) != ASTRO_OK) return st;
    GET(id) GET(flags) GET(death_cause) GET(x) GET(y) GET(z) GET(vx) GET(vy) GET(vz) GET(energy) GET(temp_cell)
    GET(emit_power) GET(dir_x) GET(dir_y) GET(dir_z) GET(biomass) GET(co2_held) GET(age_s) GET(rng_state)
    GET(taxis_memory) GET(run_timer) GET(t_local) GET(irradiance) GET(co2_local) GET(n2_local)
#undef GET
    return st;
}

uint64_t fnv1a64(const void* data, size_t n, uint64_t h) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

// Rewrites the SoA in reference-slot order (physical == logical, no
// tombstones), by a host round trip. Used by the rare host-side edits that
// address slots directly; updates `c` (the host copy of the counters).
int32_t normalize(astro_world* w, Counters& c) {
    const int32_t n = c.count;
    std::vector<uint64_t> id(n), rng(n);
    std::vector<uint32_t> flags(n);
    std::vector<uint8_t> dc(n);
    std::vector<double> x(n), y(n), z(n), vx(n), vy(n), vz(n), energy(n), biomass(n), co2(n);
    std::vector<float> tc(n), ep(n), dxv(n), dyv(n), dzv(n), age(n), tm(n), rt(n), tl(n), ir(n), cl(n), nl(n);
    astro_cell_soa s{id.data(), flags.data(), dc.data(), x.data(), y.data(), z.data(), vx.data(), vy.data(),
                     vz.data(), energy.data(), tc.data(), ep.data(), dxv.data(), dyv.data(), dzv.data(),
                     biomass.data(), co2.data(), age.data(), rng.data(), tm.data(), rt.data(), tl.data(),
                     ir.data(), cl.data(), nl.data()};
    int32_t st = w->fp32 ? download_cells(w, w->c32[0], &s, n) : download_cells(w, w->c64[0], &s, n);
    if (st != ASTRO_OK) return st;
    st = w->fp32 ? upload_cells(w, w->c32[0], &s, n) : upload_cells(w, w->c64[0], &s, n);
    if (st != ASTRO_OK) return st;
    c.nphys = n;
    w->hash_valid = false;
    return write_counters(w, c);
}

int32_t ensure_hash(astro_world* w) {
    if (w->hash_valid) return ASTRO_OK;
    const Launch L = launch_of(w);
    if (w->fp32) launch_hash_build<float>(L, one(make_view<float>(w)), false);
    else launch_hash_build<double>(L, one(make_view<double>(w)), false);
    w->hash_valid = true;
    return hip_check(w, hipGetLastError(), "hash_build");
}

// ---------------------------------------------------------------------------
// Snapshots: the contract's file layout (contracts/snapshot_v1.h:14-57) --
// a 256-byte header, then the body: parameter overrides, the cell SoA
// [0, count) array by array in contract order, the (empty) Taumoeba arrays,
// and the T, CO2, N2 values. The body's FNV-1a 64 is the state hash.
// ---------------------------------------------------------------------------
// The contract's FNV starting value (contracts/snapshot_v1.h:59); the
// published offset basis has one more digit (14695981039346656037).
constexpr uint64_t FNV_OFFSET = 1469598103934665603ull;
constexpr uint32_t SNAP_MAGIC = 0x48505341u;   // 'ASPH'
constexpr uint32_t SNAP_VERSION = 1u;

struct SnapHeader {
    uint32_t magic, version;
    uint64_t seed, tick;
    double sim_time_s;
    int32_t cell_count, cell_capacity, taumoeba_count, taumoeba_capacity;
    int32_t n_temp, n_co2, n_n2;
    double chamber_w, chamber_h, chamber_d, physics_rate, biology_rate;
    uint64_t next_cell_id;
    int32_t override_count;
    uint8_t non_canon_run, pad[3];
    char scenario_id[64];
    char build_describe[64];
    uint64_t state_hash;
};
static_assert(sizeof(SnapHeader) == 256, "snapshot header layout");
static_assert(offsetof(SnapHeader, chamber_w) == 64 && offsetof(SnapHeader, state_hash) == 248, "snapshot header");

struct SnapOverride {
    char key[48];
    double value;
};
static_assert(sizeof(SnapOverride) == 56, "snapshot override layout");

// The overrides of the canon values a world's description carries.
std::vector<SnapOverride> overrides_of(const astro_world_desc& d) {
    std::vector<SnapOverride> out;
    auto add = [&](const char* key, double v, double canon) {
        if (v == canon) return;
        SnapOverride o;
        std::memset(&o, 0, sizeof(o));   // the key's padding is hashed
        std::strncpy(o.key, key, sizeof(o.key) - 1);
        o.value = v;
        out.push_back(o);
    };
    // the reference's key order (src/sim/snapshot.cu:79-83)
    add("PETROVA_MAX_POWER", d.petrova_max_power, ASTRO_PETROVA_MAX_POWER);
    add("PETROVA_FLASH_POWER", d.petrova_flash_power, ASTRO_PETROVA_FLASH_POWER);
    add("CO2_MASS_PER_DIVISION", d.co2_mass_per_division, ASTRO_CO2_MASS_PER_DIVISION);
    return out;
}

// Snapshot size field of one grid: n for the reference's square 2D grid,
// minus the voxel count for a 3D grid (no snapshot_v1 form), 0 when absent.
int32_t snap_grid_n(const FieldHost& g) {
    if (!g.on) return 0;
    if (g.nz == 1 && g.nx == g.ny) return g.nx;
    return -static_cast<int32_t>(g.count());
}

// Host buffers of the contract SoA for n cells.
struct HostSoa {
    std::vector<uint64_t> id, rng;
    std::vector<uint32_t> flags;
    std::vector<uint8_t> dc;
    std::vector<double> x, y, z, vx, vy, vz, energy, biomass, co2;
    std::vector<float> tc, ep, dxv, dyv, dzv, age, tm, rt, tl, ir, cl, nl;
    explicit HostSoa(size_t n)
        : id(n), rng(n), flags(n), dc(n), x(n), y(n), z(n), vx(n), vy(n), vz(n), energy(n), biomass(n), co2(n),
          tc(n), ep(n), dxv(n), dyv(n), dzv(n), age(n), tm(n), rt(n), tl(n), ir(n), cl(n), nl(n) {}
    astro_cell_soa view() {
        return astro_cell_soa{id.data(), flags.data(), dc.data(), x.data(), y.data(), z.data(), vx.data(), vy.data(),
                              vz.data(), energy.data(), tc.data(), ep.data(), dxv.data(), dyv.data(), dzv.data(),
                              biomass.data(), co2.data(), age.data(), rng.data(), tm.data(), rt.data(), tl.data(),
                              ir.data(), cl.data(), nl.data()};
    }
    // visits the arrays in contract order (contracts/cell_store_v1.h:42-78)
    template <class F>
    void each(F&& f) {
        f(id); f(flags); f(dc); f(x); f(y); f(z); f(vx); f(vy); f(vz); f(energy); f(tc); f(ep); f(dxv); f(dyv);
        f(dzv); f(biomass); f(co2); f(age); f(rng); f(tm); f(rt); f(tl); f(ir); f(cl); f(nl);
    }
};

// Host buffers of the Taumoeba SoA for n predators.
struct HostTau {
    std::vector<uint64_t> id, rng;
    std::vector<uint32_t> flags, gen;
    std::vector<double> x, y, z, vx, vy, vz, biomass, prey;
    std::vector<float> dx, dy, dz, digest, tol, ema, run;
    std::vector<int32_t> target;
    explicit HostTau(size_t n)
        : id(n), rng(n), flags(n), gen(n), x(n), y(n), z(n), vx(n), vy(n), vz(n), biomass(n), prey(n), dx(n), dy(n),
          dz(n), digest(n), tol(n), ema(n), run(n), target(n) {}
    astro_tau_soa view() {
        return astro_tau_soa{id.data(),      flags.data(), x.data(),      y.data(),      z.data(),
                             vx.data(),      vy.data(),    vz.data(),     dx.data(),     dy.data(),
                             dz.data(),      biomass.data(), prey.data(), digest.data(), tol.data(),
                             gen.data(),     rng.data(),   ema.data(),    run.data(),    target.data()};
    }
    // the reference's serialization order (src/sim/snapshot.cu:59-67)
    template <class F>
    void each(F&& f) {
        f(id); f(flags); f(x); f(y); f(z); f(vx); f(vy); f(vz); f(dx); f(dy); f(dz); f(biomass); f(prey);
        f(digest); f(tol); f(gen); f(rng); f(ema); f(run); f(target);
    }
};

// The store's counters, read and written in the world's stream order.
int32_t read_tau(astro_world* w, TauCtr& tc) {
    int32_t st = hip_check(w, hipMemcpyAsync(&tc, w->tau.ctr, sizeof(tc), hipMemcpyDeviceToHost, w->stream),
                           "read taumoeba counters");
    if (st != ASTRO_OK) return st;
    return hip_check(w, hipStreamSynchronize(w->stream), "synchronize");
}

int32_t write_tau(astro_world* w, const TauCtr& tc) {
    int32_t st = hip_check(w, hipMemcpyAsync(w->tau.ctr, &tc, sizeof(tc), hipMemcpyHostToDevice, w->stream),
                           "write taumoeba counters");
    if (st != ASTRO_OK) return st;
    return hip_check(w, hipStreamSynchronize(w->stream), "synchronize");
}

// Device <-> host copies of the Taumoeba arrays [0, n), in one order.
int32_t copy_tau(astro_world* w, const astro_tau_soa* h, int32_t n, bool to_device) {
    const TauDev& t = w->tau;
    const size_t m = static_cast<size_t>(n);
    const std::pair<void*, void*> pairs[] = {
        {t.id, h->id}, {t.flags, h->flags}, {t.x, h->x}, {t.y, h->y}, {t.z, h->z}, {t.vx, h->vx},
        {t.vy, h->vy}, {t.vz, h->vz}, {t.dir_x, h->dir_x}, {t.dir_y, h->dir_y}, {t.dir_z, h->dir_z},
        {t.biomass, h->biomass}, {t.prey_biomass, h->prey_biomass}, {t.digest_timer, h->digest_timer},
        {t.tolerance, h->tolerance}, {t.generation, h->generation}, {t.rng_state, h->rng_state},
        {t.density_ema, h->density_ema}, {t.run_timer, h->run_timer}, {t.target, h->target}};
    const size_t bytes[] = {8, 4, 8, 8, 8, 8, 8, 8, 4, 4, 4, 8, 8, 4, 4, 4, 8, 4, 4, 4};
    for (size_t k = 0; k < sizeof(bytes) / sizeof(bytes[0]); ++k)
        if (!pairs[k].second) return ASTRO_INVALID_ARGUMENT;
    int32_t st = hip_check(w, hipStreamSynchronize(w->stream), "synchronize");
    for (size_t k = 0; k < sizeof(bytes) / sizeof(bytes[0]) && st == ASTRO_OK; ++k) {
        const hipError_t e =
            to_device ? hipMemcpyAsync(pairs[k].first, pairs[k].second, bytes[k] * m, hipMemcpyHostToDevice, w->stream)
                      : hipMemcpyAsync(pairs[k].second, pairs[k].first, bytes[k] * m, hipMemcpyDeviceToHost, w->stream);
        st = hip_check(w, e, "taumoeba copy");
    }
    if (st != ASTRO_OK) return st;
    return hip_check(w, hipStreamSynchronize(w->stream), "synchronize");
}

// Everything after the header, in file order (src/sim/snapshot.cu:96-118).
int32_t serialize_body(astro_world* w, std::vector<char>& body, int32_t& override_count, int32_t& count) {
    body.clear();
    Counters c{};
    int32_t st = read_counters(w, c);
    if (st != ASTRO_OK) return st;
    count = c.count;
    const std::vector<SnapOverride> ov = overrides_of(w->desc);
    override_count = static_cast<int32_t>(ov.size());
    auto put = [&](const void* p, size_t bytes) {
        const char* q = static_cast<const char*>(p);
        body.insert(body.end(), q, q + bytes);
    };
    if (!ov.empty()) put(ov.data(), sizeof(SnapOverride) * ov.size());
    if (count > 0) {
        HostSoa h(count);
        astro_cell_soa v = h.view();
        st = w->fp32 ? download_cells(w, w->c32[0], &v, count) : download_cells(w, w->c64[0], &v, count);
        if (st != ASTRO_OK) return st;
        h.each([&](const auto& a) { put(a.data(), sizeof(a[0]) * a.size()); });
    }
    TauCtr tc{};
    if ((st = read_tau(w, tc)) != ASTRO_OK) return st;
    if (tc.count > 0) {
        HostTau h(static_cast<size_t>(tc.count));
        astro_tau_soa v = h.view();
        if ((st = copy_tau(w, &v, tc.count, false)) != ASTRO_OK) return st;
        h.each([&](const auto& a) { put(a.data(), sizeof(a[0]) * a.size()); });
    }
    for (int fi : {ASTRO_FIELD_TEMPERATURE, ASTRO_FIELD_CO2, ASTRO_FIELD_N2}) {
        const FieldHost& g = w->f[fi];
        if (!g.on) continue;
        std::vector<float> v(g.count());
        if ((st = hip_check(w, hipMemcpy(v.data(), g.value, sizeof(float) * v.size(), hipMemcpyDeviceToHost),
                            "snapshot field")) != ASTRO_OK)
            return st;
        put(v.data(), sizeof(float) * v.size());
    }
    return ASTRO_OK;
}

// ---------------------------------------------------------------------------
// Ensembles: independent worlds (chambers) stepped in lock step, each stage one
// batched launch over every chamber, a whole step one graph launch.
// ---------------------------------------------------------------------------

// What the host decides for a tick; every chamber of an ensemble must agree.
struct TickShape {
    bool fp32;
    bool philox;
    double dt;
    bool thermal, light, taxis, lifecycle, contact, predators, flash;
    bool on[ASTRO_FIELD_COUNT];
    int substeps[3];   // T (when thermal), CO2, N2 (when on)
    bool operator==(const TickShape&) const = default;
};

TickShape tick_shape(const astro_world* w) {
    TickShape s{};
    const astro_world_desc& d = w->desc;
    s.fp32 = w->fp32;
    s.philox = d.rng_kind == ASTRO_RNG_PHILOX;
    s.dt = ASTRO_DT_PHYSICS * w->physics_rate;
    s.thermal = d.thermal_enabled != 0;
    s.light = d.emission_enabled && lit(w);
    s.taxis = d.taxis_enabled != 0;
    s.lifecycle = d.lifecycle_enabled != 0;
    s.contact = d.contact_enabled != 0;
    s.predators = w->tau_on;
    s.flash = w->flash_on;
    for (int i = 0; i < ASTRO_FIELD_COUNT; ++i) s.on[i] = w->f[i].on;
    if (s.thermal) s.substeps[0] = substeps_for_dt(w->f[ASTRO_FIELD_TEMPERATURE], s.dt);
    if (s.on[ASTRO_FIELD_CO2]) s.substeps[1] = substeps_for_dt(w->f[ASTRO_FIELD_CO2], s.dt);
    if (s.on[ASTRO_FIELD_N2]) s.substeps[2] = substeps_for_dt(w->f[ASTRO_FIELD_N2], s.dt);
    return s;
}

}  // namespace

struct astro_ensemble {
    std::vector<astro_world*> worlds;   // borrowed; the caller owns them
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;                // end of the last step, waited on by every chamber
    std::vector<hipEvent_t> ready;            // per chamber: its own stream's pending work

    // The captured step: `graph_ticks` ticks whose launches read their
    // arguments from `arena` (device memory written once at capture).
    char* arena = nullptr;
    size_t arena_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_ticks = 0;
    // Each chamber's generation when the graph was captured: a chamber whose
    // generation has moved since (its light, clock, flash or field buffers
    // changed) has launch arguments the graph no longer holds.
    std::vector<uint64_t> captured_gen;

    // Eagerly issued ticks (profiling, or a tick left over from the graph's
    // period) stage their arguments here.
    char* eager = nullptr;
    size_t eager_bytes = 0;

    // Blocks a grid-stride stage is spread over, across all chambers, by
    // default and per stage (set by astro_ensemble_autotune), and the FTCS
    // row-group widths per field stage.
    int stride_blocks = 2048;
    std::map<std::string, int> stage_blocks;
    std::map<std::string, int> ftcs_lanes;

    std::unique_ptr<EventProfiler> prof;
    bool prof_on = false;
    std::string err;
};

namespace {

int32_t efail(astro_ensemble* e, int32_t st, const std::string& what) {
    e->err = what;
    return st;
}

int32_t echeck(astro_ensemble* e, hipError_t r, const char* what) {
    if (r == hipSuccess) return ASTRO_OK;
    return efail(e, r == hipErrorOutOfMemory ? ASTRO_OUT_OF_MEMORY : ASTRO_DEVICE_ERROR,
                 std::string(what) + ": " + hipGetErrorString(r));
}

void drop_ensemble_graph(astro_ensemble* e) {
    if (e->gexec) hipGraphExecDestroy(e->gexec);
    if (e->graph) hipGraphDestroy(e->graph);
    if (e->arena) hipFree(e->arena);
    e->gexec = nullptr;
    e->graph = nullptr;
    e->arena = nullptr;
    e->arena_bytes = 0;
    e->graph_ticks = 0;
}

// Records `ticks` ticks of every chamber. The worlds' host state (field
// ping-pong, hash validity) advances as the launches are recorded.
int32_t record_ticks(astro_ensemble* e, Recorder& rec, int ticks) {
    const WorldSpan ws(e->worlds.data(), e->worlds.size());
    rec.stride_blocks = e->stride_blocks;
    rec.stage_blocks = e->stage_blocks;
    rec.ftcs_lanes = e->ftcs_lanes;
    const Launch L{e->stream, nullptr, &rec};
    for (int i = 0; i < ticks; ++i) {
        const int32_t st = tick_once(ws, L);
        if (st != ASTRO_OK) {
            for (astro_world* w : e->worlds)
                if (!w->err.empty()) return efail(e, st, w->err);
            return efail(e, st, "tick");
        }
    }
    return ASTRO_OK;
}

// Copies the recorded arguments into device memory `dst` (grown to fit).
int32_t upload_staging(astro_ensemble* e, const Recorder& rec, char*& dst, size_t& bytes) {
    const size_t need = std::max<size_t>(rec.staging.size(), 256);
    if (bytes < need) {
        int32_t st = echeck(e, hipStreamSynchronize(e->stream), "synchronize");
        if (st != ASTRO_OK) return st;
        if (dst) hipFree(dst);
        dst = nullptr;
        bytes = 0;
        if ((st = echeck(e, hipMalloc(&dst, need), "ensemble arguments")) != ASTRO_OK) return st;
        bytes = need;
    }
    // The previous use of the buffer is ordered before this copy on the stream,
    // but the host source must outlive the copy: synchronize around it.
    int32_t st = echeck(e, hipStreamSynchronize(e->stream), "synchronize");
    if (st != ASTRO_OK) return st;
    return echeck(e, hipMemcpy(dst, rec.staging.data(), rec.staging.size(), hipMemcpyHostToDevice), "upload arguments");
}

int32_t capture_step(astro_ensemble* e) {
    drop_ensemble_graph(e);
    const int k = ticks_per_graph(e->worlds[0]);
    std::vector<bool> hv;
    for (astro_world* w : e->worlds) hv.push_back(w->hash_valid);
    Recorder rec;
    int32_t st = record_ticks(e, rec, k);
    for (size_t i = 0; i < e->worlds.size(); ++i) e->worlds[i]->hash_valid = hv[i];
    if (st != ASTRO_OK) return st;
    if ((st = upload_staging(e, rec, e->arena, e->arena_bytes)) != ASTRO_OK) return st;
    if ((st = echeck(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal), "begin capture")) != ASTRO_OK)
        return st;
    for (const Recorder::Entry& r : rec.launches) r.fn(e->stream, e->arena);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(e->stream, &g);
    if ((st = echeck(e, ee, "end capture")) != ASTRO_OK) {
        if (g) hipGraphDestroy(g);
        return st;
    }
    e->graph = g;
    if ((st = echeck(e, hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0), "instantiate")) != ASTRO_OK) {
        drop_ensemble_graph(e);
        return st;
    }
    e->graph_ticks = k;
    e->captured_gen.clear();
    for (astro_world* w : e->worlds) e->captured_gen.push_back(w->generation);
    return ASTRO_OK;
}

// True when some chamber changed in a way that dropped its own captured tick
// since the ensemble's graph was captured.
bool chambers_changed(const astro_ensemble* e) {
    if (e->captured_gen.size() != e->worlds.size()) return true;
    for (size_t i = 0; i < e->worlds.size(); ++i)
        if (e->worlds[i]->generation != e->captured_gen[i]) return true;
    return false;
}

int32_t eager_tick(astro_ensemble* e) {
    Recorder rec;
    int32_t st = record_ticks(e, rec, 1);
    if (st != ASTRO_OK) return st;
    if ((st = upload_staging(e, rec, e->eager, e->eager_bytes)) != ASTRO_OK) return st;
    Profiler* prof = e->prof_on ? e->prof.get() : nullptr;
    for (const Recorder::Entry& r : rec.launches) {
        if (prof) prof->begin(r.name);
        r.fn(e->stream, e->eager);
        if (prof) prof->end();
    }
    return echeck(e, hipGetLastError(), "ensemble tick");
}

int32_t check_shape(astro_ensemble* e) {
    const TickShape s0 = tick_shape(e->worlds[0]);
    for (size_t i = 0; i < e->worlds.size(); ++i)
        if (e->worlds[i]->tau_on)
            return efail(e, ASTRO_UNSUPPORTED, "chamber " + std::to_string(i) + " holds Taumoeba; predation is "
                         "stepped per world, not batched");
    for (size_t i = 1; i < e->worlds.size(); ++i)
        if (!(tick_shape(e->worlds[i]) == s0))
            return efail(e, ASTRO_INVALID_ARGUMENT,
                         "chamber " + std::to_string(i) + " differs from chamber 0 in precision, clock, enabled "
                         "stages, fields or substep counts");
    return ASTRO_OK;
}

// Captures `ticks` ticks of w starting from the buffers `from` into
// graph[p][ticks - 1], and returns the buffers they leave in *to. The host's
// own buffers and hash flag are left as they were.
int32_t capture_ticks(astro_world* w, int p, int ticks, const BufPtrs& from, BufPtrs* to) {
    const BufPtrs keep = bufs_of(w);
    const bool hv = w->hash_valid;
    set_bufs(w, from);
    int32_t st = hip_check(w, hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    if (st != ASTRO_OK) {
        set_bufs(w, keep);
        return st;
    }
    int32_t cst = ASTRO_OK;
    for (int i = 0; i < ticks && cst == ASTRO_OK; ++i) cst = tick_once(w, Launch{w->stream, nullptr, nullptr});
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(w->stream, &g);
    *to = bufs_of(w);
    set_bufs(w, keep);
    w->hash_valid = hv;
    if (cst != ASTRO_OK) {
        if (g) hipGraphDestroy(g);
        return cst;
    }
    if ((st = hip_check(w, ee, "end capture")) != ASTRO_OK) return st;
    w->graph[p][ticks - 1] = g;
    if ((st = hip_check(w, hipGraphInstantiate(&w->gexec[p][ticks - 1], g, nullptr, nullptr, 0), "instantiate")) !=
        ASTRO_OK)
        return st;
    // uploaded now, so that its first launch carries no set-up
    return hip_check(w, hipGraphUpload(w->gexec[p][ticks - 1], w->stream), "graph upload");
}

// Captures the world's tick graphs from its current buffers (phase 0): one
// tick, and, when some field's substep count is odd, two ticks from phase 0
// and one and two ticks from phase 1.
int32_t capture_graphs(astro_world* w) {
    drop_graph(w);
    const bool odd = ticks_per_graph(w) == 2;
    BufPtrs end{};
    w->gstart[0] = bufs_of(w);
    int32_t st = capture_ticks(w, 0, 1, w->gstart[0], &w->gstart[1]);
    if (st == ASTRO_OK && odd) st = capture_ticks(w, 0, 2, w->gstart[0], &end);
    if (st == ASTRO_OK && odd) st = capture_ticks(w, 1, 1, w->gstart[1], &end);
    if (st == ASTRO_OK && odd) st = capture_ticks(w, 1, 2, w->gstart[1], &end);
    if (st != ASTRO_OK) {
        drop_graph(w);
        return st;
    }
    w->graph_ok = true;
    return ASTRO_OK;
}

}  // namespace


// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void astro_world_desc_default(astro_world_desc* d) {
    std::memset(d, 0, sizeof(*d));
    d->abi_version = ASTRO_ABI_VERSION;
    d->seed = 20260802ull;
    d->chamber_w = ASTRO_CHAMBER_W;
    d->chamber_h = ASTRO_CHAMBER_H;
    d->chamber_d = ASTRO_CHAMBER_D;
    d->precision = ASTRO_FP64_CONTRACT;
    d->boundary_x = d->boundary_y = ASTRO_BOUNDARY_REFLECTING;
    d->gravity_axis = ASTRO_GRAVITY_Y;
    d->thermal_noise = d->contact_enabled = d->adhesion_enabled = 1;
    d->thermal_enabled = d->emission_enabled = d->taxis_enabled = d->occlusion_exact = 1;
    d->store_disposition = ASTRO_STORE_VOID;
    d->lifecycle_enabled = 1;
    d->ambient_temp = ASTRO_AMBIENT_TEMP_DEFAULT;
    d->physics_rate = 1.0;
    d->biology_rate = 1.0;
    d->petrova_max_power = ASTRO_PETROVA_MAX_POWER;
    d->co2_mass_per_division = ASTRO_CO2_MASS_PER_DIVISION;
    d->petrova_flash_power = ASTRO_PETROVA_FLASH_POWER;
    d->light_dir_x = 1.0f;
}

const char* astro_status_str(int32_t s) {
    static const char* names[] = {"Ok", "OutOfMemory", "DeviceError", "InvalidArgument", "NotFound",
                                  "ParseError", "ContractVersionMismatch", "CapacityExceeded",
                                  "FileIoError", "Unsupported"};
    return (s >= 0 && s <= 9) ? names[s] : "Unknown";
}

const char* astro_last_error(const astro_world* w) { return w ? w->err.c_str() : ""; }

int32_t astro_world_desc_get(const astro_world* w, astro_world_desc* out) {
    if (!w || !out) return ASTRO_INVALID_ARGUMENT;
    *out = w->desc;
    return ASTRO_OK;
}

void astro_world_destroy(astro_world* w) {
    if (!w) return;
    if (w->stream) hipStreamSynchronize(w->stream);
    drop_graph(w);
    w->prof.reset();
    if (w->pool) hipFree(w->pool);
    if (w->stream) hipStreamDestroy(w->stream);
    delete w;
}

int32_t astro_world_create(const astro_world_desc* din, astro_world** out) {
    if (!din || !out) return ASTRO_INVALID_ARGUMENT;
    *out = nullptr;
    if (din->abi_version != ASTRO_ABI_VERSION) return ASTRO_CONTRACT_VERSION_MISMATCH;
    astro_world_desc d = *din;
    const int32_t cap = d.capacity > 0 ? d.capacity : ASTRO_DEFAULT_CELLS;
    if (cap <= 0 || cap > ASTRO_MAX_CELLS) return ASTRO_INVALID_ARGUMENT;
    if (!(d.chamber_w > 0 && d.chamber_h > 0 && d.chamber_d > 0)) return ASTRO_INVALID_ARGUMENT;
    if (d.precision != ASTRO_FP64_CONTRACT && d.precision != ASTRO_FP32_HOT) return ASTRO_INVALID_ARGUMENT;
    if (d.rng_kind != ASTRO_RNG_PCG32 && d.rng_kind != ASTRO_RNG_PHILOX) return ASTRO_INVALID_ARGUMENT;
    if (!(d.petrova_flash_power >= 0.0)) return ASTRO_INVALID_ARGUMENT;
    d.capacity = cap;
    const int32_t tau_cap = d.tau_capacity > 0 ? d.tau_capacity : ASTRO_DEFAULT_TAUMOEBA;
    if (tau_cap > ASTRO_MAX_TAUMOEBA) return ASTRO_INVALID_ARGUMENT;
    d.tau_capacity = tau_cap;
    if (d.division_offset <= 0.0) d.division_offset = ASTRO_CELL_RADIUS;
    const int defaults[4] = {ASTRO_FIELD_N_TEMP, ASTRO_FIELD_N_CO2, ASTRO_FIELD_N_N2, ASTRO_FIELD_N_IRRAD};
    for (int f = 0; f < 4; ++f) {
        astro_grid_dims& g = d.grid[f];
        if (g.nx == 0 && g.ny == 0 && g.nz == 0) { g.nx = g.ny = defaults[f]; g.nz = 1; }
        if (f == ASTRO_FIELD_N2 && g.nx < 0) continue;   // disabled
        if (g.nx <= 1 || g.ny <= 1 || g.nz < 1) return ASTRO_INVALID_ARGUMENT;
        // voxel indices travel as uint32 with the top bit as a flag
        if (static_cast<int64_t>(g.nx) * g.ny * g.nz >= 0x7FFFFFFFll) return ASTRO_INVALID_ARGUMENT;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ASTRO_DEVICE_ERROR;

    auto* w = new astro_world();
    w->desc = d;
    w->capacity = cap;
    w->pcap = 2 * cap;
    w->fp32 = d.precision == ASTRO_FP32_HOT;
    w->physics_rate = d.physics_rate > 0 ? d.physics_rate : 1.0;
    w->biology_rate = d.biology_rate > 0 ? d.biology_rate : 1.0;
    hipGetDevice(&w->device);
    if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) {
        delete w;
        return ASTRO_DEVICE_ERROR;
    }
    const double W = d.chamber_w, H = d.chamber_h, Dd = d.chamber_d;
    field_init(w->f[ASTRO_FIELD_TEMPERATURE], d.grid[0], W, H, Dd, ASTRO_WATER_THERMAL_DIFFUSIVITY, d.ambient_temp,
               ASTRO_DEPOSIT_SCALE_TEMPERATURE, ASTRO_BC_ROBIN);
    field_init(w->f[ASTRO_FIELD_CO2], d.grid[1], W, H, Dd, ASTRO_CO2_DIFFUSIVITY_WATER, d.co2_init,
               ASTRO_DEPOSIT_SCALE_CO2, ASTRO_BC_NEUMANN);
    field_init(w->f[ASTRO_FIELD_N2], d.grid[2], W, H, Dd, ASTRO_N2_DIFFUSIVITY, 0.0, ASTRO_DEPOSIT_SCALE_N2,
               ASTRO_BC_NEUMANN);
    field_init(w->f[ASTRO_FIELD_IRRADIANCE], d.grid[3], W, H, Dd, ASTRO_IRRADIANCE_NOMINAL_DIFFUSIVITY, 0.0,
               ASTRO_DEPOSIT_SCALE_IRRADIANCE, ASTRO_BC_NEUMANN);
    // hash_create (src/sim/hash.cu:46-80), hash_view (:90-104)
    w->hcell = d.hash_cell_size > 0.0 ? d.hash_cell_size : ASTRO_HASH_CELL_FACTOR * ASTRO_CELL_DIAMETER;
    w->hg.nx = std::max(1, static_cast<int32_t>(std::ceil(W / w->hcell)));
    w->hg.ny = std::max(1, static_cast<int32_t>(std::ceil(H / w->hcell)));
    w->hg.nz = std::max(1, static_cast<int32_t>(std::ceil(Dd / w->hcell)));
    const int64_t nb64 = static_cast<int64_t>(w->hg.nx) * w->hg.ny * w->hg.nz;
    if (nb64 >= (1ll << 30)) {
        delete w;
        return ASTRO_INVALID_ARGUMENT;
    }
    w->hg.nb = static_cast<int32_t>(nb64);
    w->hg.inv = static_cast<float>(1.0 / w->hcell);
    w->hg.ox = static_cast<float>(-0.5 * w->hg.nx * w->hcell);
    w->hg.oy = static_cast<float>(-0.5 * w->hg.ny * w->hcell);
    w->hg.oz = static_cast<float>(-0.5 * w->hg.nz * w->hcell);

    // Everything in one allocation, carved twice (measure, then assign).
    const size_t n = static_cast<size_t>(cap);
    const size_t nb = static_cast<size_t>(w->hg.nb);
    const size_t rec_bytes = w->fp32 ? sizeof(Rec<float>) : sizeof(Rec<double>);
    const size_t pn = 2 * n;
    auto layout = [&](Carver& cv) {
        if (w->fp32) {
            carve_cells(cv, w->c32[0], pn);
            carve_cells(cv, w->c32[1], n);
        } else {
            carve_cells(cv, w->c64[0], pn);
            carve_cells(cv, w->c64[1], n);
        }
        w->ctr = cv.take<Counters>(1);
        for (int fi = 0; fi < ASTRO_FIELD_COUNT; ++fi) {
            FieldHost& g = w->f[fi];
            const size_t cnt = g.count();
            g.value = cnt ? cv.take<float>(cnt) : nullptr;
            g.scratch = cnt ? cv.take<float>(cnt) : nullptr;
            g.deposit = cnt ? cv.take<unsigned long long>(cnt) : nullptr;
        }
        w->keys = cv.take<uint32_t>(pn);
        w->rin = cv.take<uint32_t>(pn);
        w->co2_avail = cv.take<float>(pn);
        w->bcount = cv.take<uint32_t>(nb + 1);
        w->bstart = cv.take<int32_t>(nb + 2);
        w->trec = cv.take<char>(n * rec_bytes);
        w->tphys = cv.take<uint32_t>(n);
        w->rec = cv.take<char>(n * rec_bytes);
        w->vals_sorted = cv.take<uint32_t>(n);
        w->sphys = cv.take<uint32_t>(n);
        w->spos = cv.take<int32_t>(pn);
        w->fs = cv.take<char>(n * (w->fp32 ? sizeof(Force<float>) : sizeof(Force<double>)));
        w->dirs = cv.take<double>(n);
        w->dflag = cv.take<int32_t>(n);
        w->prefix = cv.take<int32_t>(n);
        w->drop = cv.take<int32_t>(n);
        w->dprefix = cv.take<int32_t>(n);
        w->dlist = cv.take<uint32_t>(n);
        w->bsum = cv.take<int32_t>(static_cast<size_t>(scan_tiles_for(static_cast<int>(nb + 1))));
        w->dsum = cv.take<int32_t>(static_cast<size_t>(scan_tiles_for(cap)));
        w->xsum = cv.take<int32_t>(static_cast<size_t>(scan_tiles_for(cap)));
        w->co2_demand = cv.take<uint32_t>(std::max<size_t>(1, w->f[ASTRO_FIELD_CO2].count()));   // zeroed with the pool
        w->tcount = cv.take<uint32_t>(std::max<size_t>(1, w->f[ASTRO_FIELD_TEMPERATURE].count()));   // zeroed with the pool
        w->talone = cv.take<uint8_t>(pn);
        w->stats = cv.take<unsigned long long>(SW_COUNT);
        w->flash_accum = cv.take<unsigned long long>(4);   // zeroed with the pool
        TauDev& t = w->tau;
        const size_t tn = static_cast<size_t>(tau_cap);
        t.capacity = tau_cap;
        t.id = cv.take<uint64_t>(tn);
        t.flags = cv.take<uint32_t>(tn);
        t.x = cv.take<double>(tn);
        t.y = cv.take<double>(tn);
        t.z = cv.take<double>(tn);
        t.vx = cv.take<double>(tn);
        t.vy = cv.take<double>(tn);
        t.vz = cv.take<double>(tn);
        t.dir_x = cv.take<float>(tn);
        t.dir_y = cv.take<float>(tn);
        t.dir_z = cv.take<float>(tn);
        t.biomass = cv.take<double>(tn);
        t.prey_biomass = cv.take<double>(tn);
        t.digest_timer = cv.take<float>(tn);
        t.tolerance = cv.take<float>(tn);
        t.generation = cv.take<uint32_t>(tn);
        t.rng_state = cv.take<uint64_t>(tn);
        t.density_ema = cv.take<float>(tn);
        t.run_timer = cv.take<float>(tn);
        t.target = cv.take<int32_t>(tn);
        t.ctr = cv.take<TauCtr>(1);
        t.prey_rec = cv.take<int32_t>(tn);
        t.mark = cv.take<int32_t>(tn);
        t.prefix = cv.take<int32_t>(tn);
        t.src = cv.take<int32_t>(tn);
        t.scratch = cv.take<unsigned long long>(tn);
        t.claim = cv.take<unsigned long long>(n);
        t.cell_capacity = cap;
    };
    Carver measure;
    layout(measure);
    w->pool_bytes = measure.off;
    hipError_t e = hipMalloc(&w->pool, w->pool_bytes);
    if (e != hipSuccess) {
        hipStreamDestroy(w->stream);
        delete w;
        return e == hipErrorOutOfMemory ? ASTRO_OUT_OF_MEMORY : ASTRO_DEVICE_ERROR;
    }
    Carver assign;
    assign.base = static_cast<char*>(w->pool);
    layout(assign);
    if (hipMemsetAsync(w->pool, 0, w->pool_bytes, w->stream) != hipSuccess) {
        astro_world_destroy(w);
        return ASTRO_DEVICE_ERROR;
    }
    Counters c0{};
    c0.count = 0;
    c0.capacity = cap;
    c0.next_id = 1;
    hipMemcpyAsync(w->ctr, &c0, sizeof(c0), hipMemcpyHostToDevice, w->stream);
    w->tau_cap = tau_cap;
    TauCtr t0{};
    t0.capacity = tau_cap;
    t0.next_id = 1;
    hipMemcpyAsync(w->tau.ctr, &t0, sizeof(t0), hipMemcpyHostToDevice, w->stream);
    const Launch L{w->stream, nullptr};
    launch_fill_u64(L, w->tau.claim, n, ~0ull);   // no claims
    for (int fi = 0; fi < ASTRO_FIELD_COUNT; ++fi) {
        FieldHost& g = w->f[fi];
        launch_fill(L, g.value, g.count(), static_cast<float>(g.ambient));
        launch_fill(L, g.scratch, g.count(), static_cast<float>(g.ambient));
    }
    w->prof = std::make_unique<EventProfiler>();
    w->prof->stream = w->stream;
    if (hipStreamSynchronize(w->stream) != hipSuccess || hipGetLastError() != hipSuccess) {
        astro_world_destroy(w);
        return ASTRO_DEVICE_ERROR;
    }
    *out = w;
    return ASTRO_OK;
}

int32_t astro_cells_spawn(astro_world* w, const astro_spawn_params* p, uint64_t seed) {
    if (!w || !p) return ASTRO_INVALID_ARGUMENT;
    if (p->count <= 0) return ASTRO_OK;
    if (p->charge_dist == ASTRO_DIST_LOG_UNIFORM && !(p->charge_a > 0.0 && p->charge_b > 0.0))
        return fail(w, ASTRO_INVALID_ARGUMENT, "log-uniform charge needs positive bounds");
    Counters c{};
    int32_t st = read_counters(w, c);
    if (st != ASTRO_OK) return st;
    if (c.count + p->count > w->capacity) return fail(w, ASTRO_CAPACITY_EXCEEDED, "spawn exceeds cell store capacity");
    if (c.nphys + p->count > w->pcap) {   // pack tombstones first
        if ((st = normalize(w, c)) != ASTRO_OK) return st;
    }
    SpawnArgs a{};
    a.count = p->count; a.placement = p->placement; a.charge_dist = p->charge_dist; a.awake = p->awake;
    a.place_x = p->place_x; a.place_y = p->place_y; a.place_radius = p->place_radius;
    a.charge_a = p->charge_a; a.charge_b = p->charge_b;
    a.chamber_w = w->desc.chamber_w; a.chamber_h = w->desc.chamber_h; a.chamber_d = w->desc.chamber_d;
    a.philox = w->desc.rng_kind == ASTRO_RNG_PHILOX ? 1 : 0;
    const Launch L = launch_of(w);
    if (w->fp32) launch_spawn<float>(L, make_view<float>(w), a, c.nphys, c.count, c.next_id, seed);
    else launch_spawn<double>(L, make_view<double>(w), a, c.nphys, c.count, c.next_id, seed);
    if ((st = hip_check(w, hipGetLastError(), "spawn_kernel")) != ASTRO_OK) return st;
    c.count += p->count;
    c.nphys += p->count;
    c.next_id += static_cast<uint64_t>(p->count);
    w->hash_valid = false;
    w->live_hint += p->count;
    if (p->awake) w->awake_hint += p->count;
    return write_counters(w, c);
}

int32_t astro_cells_upload(astro_world* w, const astro_cell_soa* s, int32_t n) {
    if (!w || !s || n < 0) return ASTRO_INVALID_ARGUMENT;
    if (n > w->capacity) return fail(w, ASTRO_CAPACITY_EXCEEDED, "upload exceeds capacity");
    Counters c{};
    int32_t st = read_counters(w, c);
    if (st != ASTRO_OK) return st;
    st = w->fp32 ? upload_cells(w, w->c32[0], s, n) : upload_cells(w, w->c64[0], s, n);
    if (st != ASTRO_OK) return st;
    if (s->flags) {
        w->live_hint = w->awake_hint = 0;
        for (int32_t i = 0; i < n; ++i) {
            const uint32_t f = s->flags[i];
            if ((f & ASTRO_FLAG_OCCUPIED) && (f & ASTRO_FLAG_ALIVE)) {
                ++w->live_hint;
                if (f & ASTRO_FLAG_AWAKE) ++w->awake_hint;
            }
        }
    }
    c.count = n;
    c.nphys = n;
    w->hash_valid = false;
    return write_counters(w, c);
}

int32_t astro_cells_download(astro_world* w, astro_cell_soa* s, int32_t n, int32_t* n_out) {
    if (!w || !s) return ASTRO_INVALID_ARGUMENT;
    Counters c{};
    int32_t st = read_counters(w, c);
    if (st != ASTRO_OK) return st;
    const int32_t m = std::min(n, c.count);
    st = w->fp32 ? download_cells(w, w->c32[0], s, m) : download_cells(w, w->c64[0], s, m);
    if (st != ASTRO_OK) return st;
    if (n_out) *n_out = m;
    return ASTRO_OK;
}

int32_t astro_world_get_counters(astro_world* w, astro_counters* o) {
    if (!w || !o) return ASTRO_INVALID_ARGUMENT;
    Counters c{};
    int32_t st = read_counters(w, c);
    if (st != ASTRO_OK) return st;
    std::memset(o, 0, sizeof(*o));
    o->count = c.count;
    o->capacity = w->capacity;
    o->next_id = c.next_id;
    o->tick = w->tick;
    o->sim_time_s = w->sim_time;
    o->divisions_total = c.divisions_total;
    o->deaths_total = c.deaths_total;
    o->dead_slots_prev = c.dead_slots_prev;
    o->cell_steps = c.cell_steps;
    return ASTRO_OK;
}

int32_t astro_world_set_counters(astro_world* w, const astro_counters* in) {
    if (!w || !in) return ASTRO_INVALID_ARGUMENT;
    if (in->count < 0 || in->count > w->capacity) return ASTRO_CAPACITY_EXCEEDED;
    Counters c{};
    int32_t st = read_counters(w, c);
    if (st != ASTRO_OK) return st;
    if ((st = normalize(w, c)) != ASTRO_OK) return st;
    c.count = in->count;
    c.nphys = in->count;
    c.next_id = in->next_id;
    c.divisions_total = in->divisions_total;
    c.deaths_total = in->deaths_total;
    c.dead_slots_prev = in->dead_slots_prev;
    c.cell_steps = in->cell_steps;
    c.tick_count = in->tick;
    w->tick = in->tick;
    w->sim_time = in->sim_time_s;
    w->hash_valid = false;
    return write_counters(w, c);
}

int32_t astro_field_dims(const astro_world* w, int32_t f, astro_grid_dims* o, double* dx, double* dy, double* dz) {
    if (!w || f < 0 || f >= ASTRO_FIELD_COUNT) return ASTRO_INVALID_ARGUMENT;
    const FieldHost& g = w->f[f];
    if (o) { o->nx = g.nx; o->ny = g.ny; o->nz = g.nz; }
    if (dx) *dx = g.dx;
    if (dy) *dy = g.dy;
    if (dz) *dz = g.dz;
    return ASTRO_OK;
}

int32_t astro_field_upload(astro_world* w, int32_t f, const float* v, size_t count) {
    if (!w || f < 0 || f >= ASTRO_FIELD_COUNT || !v) return ASTRO_INVALID_ARGUMENT;
    FieldHost& g = w->f[f];
    if (!g.on || count != g.count()) return fail(w, ASTRO_INVALID_ARGUMENT, "field size mismatch");
    int32_t st = hip_check(w, hipStreamSynchronize(w->stream), "synchronize");
    if (st != ASTRO_OK) return st;
    st = hip_check(w, hipMemcpy(g.value, v, sizeof(float) * count, hipMemcpyHostToDevice), "field upload");
    if (st != ASTRO_OK) return st;
    return hip_check(w, hipMemcpy(g.scratch, v, sizeof(float) * count, hipMemcpyHostToDevice), "field upload");
}

int32_t astro_field_download(astro_world* w, int32_t f, float* v, size_t count) {
    if (!w || f < 0 || f >= ASTRO_FIELD_COUNT || !v) return ASTRO_INVALID_ARGUMENT;
    FieldHost& g = w->f[f];
    if (!g.on || count != g.count()) return fail(w, ASTRO_INVALID_ARGUMENT, "field size mismatch");
    int32_t st = hip_check(w, hipStreamSynchronize(w->stream), "synchronize");
    if (st != ASTRO_OK) return st;
    return hip_check(w, hipMemcpy(v, g.value, sizeof(float) * count, hipMemcpyDeviceToHost), "field download");
}

int32_t astro_field_deposit_download(astro_world* w, int32_t f, int64_t* v, size_t count) {
    if (!w || f < 0 || f >= ASTRO_FIELD_COUNT || !v) return ASTRO_INVALID_ARGUMENT;
    FieldHost& g = w->f[f];
    if (!g.on || count != g.count()) return fail(w, ASTRO_INVALID_ARGUMENT, "field size mismatch");
    int32_t st = hip_check(w, hipStreamSynchronize(w->stream), "synchronize");
    if (st != ASTRO_OK) return st;
    return hip_check(w, hipMemcpy(v, g.deposit, sizeof(int64_t) * count, hipMemcpyDeviceToHost), "deposit download");
}

int32_t astro_world_step(astro_world* w, int32_t n_ticks) {
    if (!w || n_ticks < 0) return ASTRO_INVALID_ARGUMENT;
    if (n_ticks == 0) return ASTRO_OK;
    int32_t st;
    int32_t done = 0;
    const bool odd = ticks_per_graph(w) == 2;
    if (!w->prof_on && (w->graph_ok || n_ticks >= (odd ? 2 : 1))) {
        if (!w->graph_ok && (st = capture_graphs(w)) != ASTRO_OK) return st;
        // Ticks in pairs where a pair's buffers end where they started, a
        // single tick for an odd remainder; no tick runs eagerly and nothing
        // is captured again.
        while (done < n_ticks) {
            const int t = odd && n_ticks - done >= 2 ? 2 : 1;
            if ((st = hip_check(w, hipGraphLaunch(w->gexec[w->phase][t - 1], w->stream), "graph launch")) != ASTRO_OK)
                return st;
            advance_clock(w, t);
            done += t;
            if (odd && t == 1) w->phase ^= 1;
            set_bufs(w, w->gstart[w->phase]);
        }
        w->hash_valid = false;
        return ASTRO_OK;
    }
    const Launch L = launch_of(w);
    while (done < n_ticks) {
        if ((st = tick_once(w, L)) != ASTRO_OK) return st;
        if ((st = hip_check(w, hipGetLastError(), "tick launch")) != ASTRO_OK) return st;
        advance_clock(w, 1);
        ++done;
        // An eager tick swaps the buffers of a field with an odd substep
        // count, as graph[phase][0] would: the captured graphs still apply,
        // from the other phase.
        if (w->graph_ok && odd) w->phase ^= 1;
    }
    return ASTRO_OK;
}

int32_t astro_world_sync(astro_world* w) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    int32_t st = hip_check(w, hipStreamSynchronize(w->stream), "synchronize");
    if (st != ASTRO_OK) return st;
    return check_device_errors(w);
}

// world_set_clock with a preset (src/sim/step.cu:190-215): the preset's rates,
// or the custom ones, clamped.
int32_t astro_world_set_clock_preset(astro_world* w, int32_t preset, double custom_physics, double custom_biology) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    double physics = custom_physics, biology = custom_biology;
    switch (preset) {
        case ASTRO_CLOCK_MOTION: physics = ASTRO_CLOCK_MOTION_PHYSICS; biology = ASTRO_CLOCK_MOTION_BIOLOGY; break;
        case ASTRO_CLOCK_METABOLIC:
            physics = ASTRO_CLOCK_METABOLIC_PHYSICS;
            biology = ASTRO_CLOCK_METABOLIC_BIOLOGY;
            break;
        case ASTRO_CLOCK_GENERATIONAL:
            physics = ASTRO_CLOCK_GENERATIONAL_PHYSICS;
            biology = ASTRO_CLOCK_GENERATIONAL_BIOLOGY;
            break;
        case ASTRO_CLOCK_CUSTOM: break;
        default:   // realtime, and any unknown preset
            physics = ASTRO_CLOCK_REALTIME_PHYSICS;
            biology = ASTRO_CLOCK_REALTIME_BIOLOGY;
            break;
    }
    return astro_world_set_clock(w, physics, biology);
}

// World::light and World::ambient_irradiance (src/sim/world.cuh). The captured
// tick is dropped only when something changed, so a scripted light level that
// holds steady costs nothing per tick.
int32_t astro_world_set_light(astro_world* w, float dir_x, float dir_y, float irradiance, float ambient,
                              int32_t enabled) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    astro_world_desc& d = w->desc;
    const uint8_t on = enabled ? 1 : 0;
    if (d.light_dir_x == dir_x && d.light_dir_y == dir_y && d.light_irradiance == irradiance &&
        d.ambient_irradiance == ambient && d.light_enabled == on)
        return ASTRO_OK;
    d.light_dir_x = dir_x;
    d.light_dir_y = dir_y;
    d.light_irradiance = irradiance;
    d.ambient_irradiance = ambient;
    d.light_enabled = on;
    drop_graph(w);
    return ASTRO_OK;
}

// world_apply_brush (src/sim/step.cu:80-89) over grid_brush
// (src/fields/diffuse.cu:81-89); enqueued on the world's stream.
int32_t astro_world_apply_brush(astro_world* w, int32_t kind, double x, double y, double radius, double strength) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    int field = -1;
    double amount = strength;
    switch (kind) {
        case ASTRO_BRUSH_HEAT: field = ASTRO_FIELD_TEMPERATURE; break;
        case ASTRO_BRUSH_CHILL: field = ASTRO_FIELD_TEMPERATURE; amount = -strength; break;
        case ASTRO_BRUSH_INJECT_CO2: field = ASTRO_FIELD_CO2; break;
        case ASTRO_BRUSH_INJECT_N2: field = ASTRO_FIELD_N2; break;
        default: return fail(w, ASTRO_INVALID_ARGUMENT, "unknown brush");
    }
    const FieldHost& g = w->f[field];
    if (!g.on) return fail(w, ASTRO_INVALID_ARGUMENT, "brush on a disabled field");
    if (radius <= 0.0) return ASTRO_OK;
    launch_brush(Launch{w->stream, nullptr}, g.grid(), static_cast<float>(x), static_cast<float>(y),
                 static_cast<float>(radius), static_cast<float>(amount));
    return hip_check(w, hipGetLastError(), "brush");
}

// grid_fill (src/fields/diffuse.cu:128-134): value and scratch.
int32_t astro_field_fill(astro_world* w, int32_t field, float value) {
    if (!w || field < 0 || field >= ASTRO_FIELD_COUNT) return ASTRO_INVALID_ARGUMENT;
    const FieldHost& g = w->f[field];
    if (!g.on) return fail(w, ASTRO_INVALID_ARGUMENT, "fill of a disabled field");
    const Launch L{w->stream, nullptr};
    launch_fill(L, g.value, g.count(), value);
    launch_fill(L, g.scratch, g.count(), value);
    return hip_check(w, hipGetLastError(), "fill");
}

// World::flash_active (src/sim/world.cuh:178), armed per tick by the driving
// script; the captured tick is dropped only when it changes.
int32_t astro_world_set_flash(astro_world* w, int32_t armed) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    const bool on = armed != 0;
    if (on != w->flash_on) {
        w->flash_on = on;
        drop_graph(w);
    }
    return ASTRO_OK;
}

// world_flash_audit (src/sim/flash.cu:69-79).
int32_t astro_world_flash_audit(astro_world* w, double* impulse_ns, double* discharged_j) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    unsigned long long raw[4] = {0, 0, 0, 0};
    int32_t st = hip_check(w, hipMemcpyAsync(raw, w->flash_accum, sizeof(raw), hipMemcpyDeviceToHost, w->stream),
                           "flash audit");
    if (st == ASTRO_OK) st = hip_check(w, hipStreamSynchronize(w->stream), "synchronize");
    if (st != ASTRO_OK) return st;
    auto from_fixed = [](unsigned long long r, double scale) { return static_cast<double>(static_cast<int64_t>(r)) / scale; };
    const double ix = from_fixed(raw[0], ASTRO_FLASH_IMPULSE_SCALE);
    const double iy = from_fixed(raw[1], ASTRO_FLASH_IMPULSE_SCALE);
    const double iz = from_fixed(raw[2], ASTRO_FLASH_IMPULSE_SCALE);
    if (impulse_ns) *impulse_ns = std::sqrt(ix * ix + iy * iy + iz * iz);
    if (discharged_j) *discharged_j = from_fixed(raw[3], ASTRO_FLASH_ENERGY_SCALE);
    return ASTRO_OK;
}

// The host-side clock, without waiting for the device.
int32_t astro_world_time(const astro_world* w, uint64_t* tick, double* sim_time_s, double* physics_rate,
                         double* biology_rate) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    if (tick) *tick = w->tick;
    if (sim_time_s) *sim_time_s = w->sim_time;
    if (physics_rate) *physics_rate = w->physics_rate;
    if (biology_rate) *biology_rate = w->biology_rate;
    return ASTRO_OK;
}

int32_t astro_world_set_clock(astro_world* w, double physics, double biology) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    w->physics_rate = physics < ASTRO_CLOCK_PHYSICS_MIN ? ASTRO_CLOCK_PHYSICS_MIN
                    : (physics > ASTRO_CLOCK_PHYSICS_MAX ? ASTRO_CLOCK_PHYSICS_MAX : physics);
    w->biology_rate = biology < ASTRO_CLOCK_BIOLOGY_MIN ? ASTRO_CLOCK_BIOLOGY_MIN
                    : (biology > ASTRO_CLOCK_BIOLOGY_MAX ? ASTRO_CLOCK_BIOLOGY_MAX : biology);
    w->desc.physics_rate = w->physics_rate;
    w->desc.biology_rate = w->biology_rate;
    drop_graph(w);
    return ASTRO_OK;
}

int32_t astro_world_set_deposit_path(astro_world* w, int32_t path) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    if (path < ASTRO_DEPOSIT_AUTO || path > ASTRO_DEPOSIT_DIRECT)
        return fail(w, ASTRO_INVALID_ARGUMENT, "unknown deposit path");
    w->deposit_path = path;
    drop_graph(w);   // a captured tick keeps the path it was captured with
    return ASTRO_OK;
}

// stats_step + world_stats (src/sim/stats.cu:79-153).
int32_t astro_world_stats(astro_world* w, astro_stats* s) {
    if (!w || !s) return ASTRO_INVALID_ARGUMENT;
    int32_t st = hip_check(w, hipMemsetAsync(w->stats, 0, sizeof(unsigned long long) * SW_COUNT, w->stream), "stats clear");
    if (st != ASTRO_OK) return st;
    const Launch L{w->stream, nullptr};
    if (w->fp32) launch_stats_cells<float>(L, make_view<float>(w));
    else launch_stats_cells<double>(L, make_view<double>(w));
    launch_stats_fields(L, w->f[0].grid(), w->f[1].grid(), w->f[2].grid(), w->f[1].on, w->f[2].on, w->stats);
    if (w->tau_on) launch_tau_stats(L, w->tau, w->stats);
    if ((st = hip_check(w, hipGetLastError(), "stats kernels")) != ASTRO_OK) return st;
    unsigned long long a[SW_COUNT];
    Counters c{};
    if ((st = read_counters(w, c)) != ASTRO_OK) return st;
    if ((st = hip_check(w, hipMemcpy(a, w->stats, sizeof(a), hipMemcpyDeviceToHost), "stats read")) != ASTRO_OK) return st;
    auto from_fixed = [](unsigned long long raw, double scale) { return static_cast<double>(static_cast<int64_t>(raw)) / scale; };
    std::memset(s, 0, sizeof(*s));
    s->tick = w->tick;
    s->sim_time_s = w->sim_time;
    s->physics_rate = w->physics_rate;
    s->biology_rate = w->biology_rate;
    const uint32_t na = static_cast<uint32_t>(a[SW_N_ALIVE]), nd = static_cast<uint32_t>(a[SW_N_DEAD]);
    w->live_hint = na;
    w->awake_hint = static_cast<int64_t>(a[SW_N_AWAKE]);
    s->n_live = static_cast<int32_t>(na);
    s->n_dead = static_cast<int32_t>(nd);
    s->n_awake = static_cast<int32_t>(a[SW_N_AWAKE]);
    s->n_dormant = static_cast<int32_t>(a[SW_N_DORMANT]);
    s->n_taumoeba = static_cast<int32_t>(a[SW_N_TAU]);
    s->mean_tau_tolerance = a[SW_N_TAU] > 0 ? from_fixed(a[SW_TAU_TOL], ASTRO_DEPOSIT_SCALE_STATS) /
                                                  static_cast<double>(a[SW_N_TAU])
                                            : 0.0;
    s->total_energy_j = from_fixed(a[SW_ENERGY], ASTRO_DEPOSIT_SCALE_STATS);
    s->mean_charge = (na + nd) > 0u ? s->total_energy_j / ((na + nd) * ASTRO_CELL_ENERGY_MAX) : 0.0;
    s->mean_temp_cell_k = na > 0u ? from_fixed(a[SW_TEMP_CELL], ASTRO_DEPOSIT_SCALE_STATS) / na : 0.0;
    const FieldHost& T = w->f[ASTRO_FIELD_TEMPERATURE];
    const double cells = static_cast<double>(T.count());
    s->mean_temp_medium_k = cells > 0.0 ? from_fixed(a[SW_TEMP_MEDIUM], ASTRO_DEPOSIT_SCALE_STATS) / cells : 0.0;
    s->max_temp_medium_k = from_fixed(a[SW_MAX_TEMP], ASTRO_DEPOSIT_SCALE_STATS);
    const FieldHost& C = w->f[ASTRO_FIELD_CO2];
    const FieldHost& N = w->f[ASTRO_FIELD_N2];
    s->co2_total_kg = from_fixed(a[SW_CO2], ASTRO_STATS_SCALE_FIELD) * (C.dx * C.dy * C.dz);
    s->n2_total_kg = from_fixed(a[SW_N2], ASTRO_STATS_SCALE_FIELD) * (N.on ? N.dx * N.dy * N.dz : 0.0);
    s->divisions_this_window = static_cast<int32_t>(c.divisions_total - w->div_reported);
    s->deaths_this_window = static_cast<int32_t>(c.deaths_total - w->deaths_reported);
    w->div_reported = c.divisions_total;
    w->deaths_reported = c.deaths_total;
    s->boil_events = s->max_temp_medium_k >= 373.14999999999998 ? 1 : 0;
    s->non_canon_run = (w->non_canon || !overrides_of(w->desc).empty()) ? 1 : 0;
    return ASTRO_OK;
}

// snapshot_state_hash (src/sim/snapshot.cu:100-129).
int32_t astro_world_state_hash(astro_world* w, uint64_t* out) {
    if (!w || !out) return ASTRO_INVALID_ARGUMENT;
    std::vector<char> body;
    int32_t oc = 0, n = 0;
    const int32_t st = serialize_body(w, body, oc, n);
    if (st != ASTRO_OK) return st;
    *out = fnv1a64(body.data(), body.size(), FNV_OFFSET);
    return ASTRO_OK;
}

// ---- stage entry points ----------------------------------------------------
int32_t astro_stage_hash_build(astro_world* w) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    w->hash_valid = false;
    return ensure_hash(w);
}

int32_t astro_hash_dims(const astro_world* w, int32_t* nx, int32_t* ny, int32_t* nz, double* cell) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    if (nx) *nx = w->hg.nx;
    if (ny) *ny = w->hg.ny;
    if (nz) *nz = w->hg.nz;
    if (cell) *cell = w->hcell;
    return ASTRO_OK;
}

int32_t astro_hash_download(astro_world* w, uint32_t* vals, int32_t n, int32_t* bs, int32_t* be, int32_t nb) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    Counters c{};
    int32_t st = read_counters(w, c);
    if (st != ASTRO_OK) return st;
    const int32_t m = std::min(n, c.count);
    if (vals && m > 0 &&
        (st = hip_check(w, hipMemcpy(vals, w->vals_sorted, sizeof(uint32_t) * m, hipMemcpyDeviceToHost), "hash vals")) != ASTRO_OK)
        return st;
    if (bs || be) {
        std::vector<int32_t> dense(static_cast<size_t>(w->hg.nb) + 2);
        if ((st = hip_check(w, hipMemcpy(dense.data(), w->bstart, sizeof(int32_t) * dense.size(), hipMemcpyDeviceToHost),
                            "hash ranges")) != ASTRO_OK)
            return st;
        const int32_t k = std::min(nb, w->hg.nb);
        for (int32_t b = 0; b < k; ++b) {
            const bool any = dense[b + 1] > dense[b];
            if (bs) bs[b] = any ? dense[b] : 0;
            if (be) be[b] = any ? dense[b + 1] : 0;
        }
    }
    return ASTRO_OK;
}

int32_t astro_stage_thermal(astro_world* w, double dt) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    const Launch L = launch_of(w);
    astro_world* const one_w[1] = {w};
    const WorldSpan ws(one_w, 1);
    const TickParams t = tick_params(w, dt);
    const int32_t st = w->fp32 ? enqueue_thermal<float>(ws, L, dt, one(t)) : enqueue_thermal<double>(ws, L, dt, one(t));
    drop_graph(w);
    if (st != ASTRO_OK) return st;
    return hip_check(w, hipGetLastError(), "thermal");
}

int32_t astro_stage_emission(astro_world* w, double dt) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    if (!lit(w)) return ASTRO_OK;
    int32_t st = ensure_hash(w);
    if (st != ASTRO_OK) return st;
    const Launch L = launch_of(w);
    const TickParams t = tick_params(w, dt);
    astro_world* const one_w[1] = {w};
    const WorldSpan ws(one_w, 1);
    if (w->fp32) {
        enqueue_light<float>(ws, L, one(t));
        launch_feed_taxis<float>(L, one(make_view<float>(w)), one(t), true, false);
    } else {
        enqueue_light<double>(ws, L, one(t));
        launch_feed_taxis<double>(L, one(make_view<double>(w)), one(t), true, false);
    }
    return hip_check(w, hipGetLastError(), "emission");
}

// A stage called on its own draws from the current tick's counter streams.
static void set_rng_tick(astro_world* w, const Launch& L) {
    launch_fill_u64(L, reinterpret_cast<unsigned long long*>(&w->ctr->rng_tick), 1,
                    static_cast<unsigned long long>(w->tick));
}

int32_t astro_stage_taxis(astro_world* w, double dt) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    const Launch L = launch_of(w);
    const TickParams t = tick_params(w, dt);
    set_rng_tick(w, L);
    if (w->fp32) launch_feed_taxis<float>(L, one(make_view<float>(w)), one(t), false, true);
    else launch_feed_taxis<double>(L, one(make_view<double>(w)), one(t), false, true);
    return hip_check(w, hipGetLastError(), "taxis");
}

int32_t astro_stage_forces(astro_world* w) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    int32_t st = ensure_hash(w);
    if (st != ASTRO_OK) return st;
    const Launch L = launch_of(w);
    const TickParams t = tick_params(w, ASTRO_DT_PHYSICS * w->physics_rate);
    if (w->fp32) launch_cell_forces<float>(L, one(make_view<float>(w)), one(t), false, false);
    else launch_cell_forces<double>(L, one(make_view<double>(w)), one(t), false, false);
    return hip_check(w, hipGetLastError(), "forces");
}

int32_t astro_forces_download(astro_world* w, double* fx, double* fy, double* fz, int32_t n) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    Counters c{};
    int32_t st = read_counters(w, c);
    if (st != ASTRO_OK) return st;
    const int32_t m = std::min(n, c.count);
    if (m <= 0) return ASTRO_OK;
    const int32_t np = c.nphys;
    std::vector<double> f(3 * static_cast<size_t>(c.count));
    std::vector<int32_t> sp(np);
    std::vector<uint32_t> ls(np);
    const uint32_t* lslot_dev = w->fp32 ? w->c32[0].lslot : w->c64[0].lslot;
    if (w->fp32) {
        std::vector<Force<float>> ff(c.count);
        if ((st = hip_check(w, hipMemcpy(ff.data(), w->fs, sizeof(Force<float>) * c.count, hipMemcpyDeviceToHost), "forces")) != ASTRO_OK) return st;
        for (int32_t q = 0; q < c.count; ++q) {
            f[3 * q] = ff[q].x; f[3 * q + 1] = ff[q].y; f[3 * q + 2] = ff[q].z;
        }
    } else if ((st = hip_check(w, hipMemcpy(f.data(), w->fs, sizeof(Force<double>) * c.count, hipMemcpyDeviceToHost), "forces")) != ASTRO_OK) {
        return st;
    }
    if ((st = hip_check(w, hipMemcpy(sp.data(), w->spos, sizeof(int32_t) * np, hipMemcpyDeviceToHost), "spos")) != ASTRO_OK) return st;
    if ((st = hip_check(w, hipMemcpy(ls.data(), lslot_dev, sizeof(uint32_t) * np, hipMemcpyDeviceToHost), "lslot")) != ASTRO_OK) return st;
    for (int32_t p = 0; p < np; ++p) {
        const uint32_t i = ls[p];
        if (i == NO_SLOT || i >= static_cast<uint32_t>(m)) continue;
        const double* F = &f[3 * static_cast<size_t>(sp[p])];
        if (fx) fx[i] = F[0];
        if (fy) fy[i] = F[1];
        if (fz) fz[i] = F[2];
    }
    return ASTRO_OK;
}

int32_t astro_stage_integrate(astro_world* w, double dt) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    const Launch L = launch_of(w);
    const TickParams t = tick_params(w, dt);
    set_rng_tick(w, L);
    if (w->fp32) launch_integrate<float>(L, one(make_view<float>(w)), one(t), false);
    else launch_integrate<double>(L, one(make_view<double>(w)), one(t), false);
    w->hash_valid = false;
    return hip_check(w, hipGetLastError(), "integrate");
}

int32_t astro_stage_motion(astro_world* w, double dt) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    int32_t st = ensure_hash(w);
    if (st != ASTRO_OK) return st;
    const Launch L = launch_of(w);
    const TickParams t = tick_params(w, dt);
    set_rng_tick(w, L);
    if (w->fp32) {
        const WorldDev<float> d = make_view<float>(w);
        launch_cell_forces<float>(L, one(d), one(t), false, false);
        launch_integrate<float>(L, one(d), one(t), false);
    } else {
        const WorldDev<double> d = make_view<double>(w);
        launch_cell_forces<double>(L, one(d), one(t), false, false);
        launch_integrate<double>(L, one(d), one(t), false);
    }
    w->hash_valid = false;
    return hip_check(w, hipGetLastError(), "motion");
}

int32_t astro_stage_slow_fields(astro_world* w, double dt) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    const Launch L = launch_of(w);
    astro_world* const one_w[1] = {w};
    const WorldSpan ws(one_w, 1);
    int32_t st = enqueue_diffuse(ws, L, ASTRO_FIELD_CO2, dt);
    if (st == ASTRO_OK) st = enqueue_diffuse(ws, L, ASTRO_FIELD_N2, dt);
    drop_graph(w);
    if (st != ASTRO_OK) return st;
    return hip_check(w, hipGetLastError(), "slow fields");
}

int32_t astro_stage_diffuse(astro_world* w, int32_t f, double dt) {
    if (!w || f < 0 || f >= ASTRO_FIELD_COUNT) return ASTRO_INVALID_ARGUMENT;
    astro_world* const one_w[1] = {w};
    const int32_t st = enqueue_diffuse(WorldSpan(one_w, 1), launch_of(w), f, dt);
    drop_graph(w);
    if (st != ASTRO_OK) return st;
    return hip_check(w, hipGetLastError(), "diffuse");
}

int32_t astro_stage_lifecycle(astro_world* w, double dt) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    const Launch L = launch_of(w);
    const TickParams t = tick_params(w, dt);
    if (w->fp32) {
        launch_life_clear<float>(L, make_view<float>(w));
        launch_lifecycle<float>(L, one(make_view<float>(w)), one(t), false);
    } else {
        launch_life_clear<double>(L, make_view<double>(w));
        launch_lifecycle<double>(L, one(make_view<double>(w)), one(t), false);
    }
    w->hash_valid = false;
    return hip_check(w, hipGetLastError(), "lifecycle");
}

int32_t astro_stage_compact(astro_world* w) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    const Launch L = launch_of(w);
    if (w->fp32) launch_compact<float>(L, make_view<float>(w));
    else launch_compact<double>(L, make_view<double>(w));
    w->hash_valid = false;
    return hip_check(w, hipGetLastError(), "compact");
}

// ---- measurement -------------------------------------------------------------
int32_t astro_profile_enable(astro_world* w, int32_t on) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    w->prof_on = on != 0;
    return ASTRO_OK;
}

int32_t astro_profile_read(astro_world* w, char* names, int32_t stride, int64_t* launches, double* total_ms,
                           int32_t max, int32_t* n_out) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    w->prof->drain();
    int32_t k = 0;
    for (const std::string& name : w->prof->order) {
        if (k >= max) break;
        const auto& v = w->prof->totals[name];
        if (names && stride > 0) {
            std::strncpy(names + static_cast<size_t>(k) * stride, name.c_str(), stride - 1);
            names[static_cast<size_t>(k) * stride + stride - 1] = '\0';
        }
        if (launches) launches[k] = v.first;
        if (total_ms) total_ms[k] = v.second;
        ++k;
    }
    if (n_out) *n_out = k;
    return ASTRO_OK;
}

int32_t astro_profile_reset(astro_world* w) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    w->prof->drain();
    w->prof->totals.clear();
    w->prof->order.clear();
    return ASTRO_OK;
}

void* astro_world_stream(astro_world* w) { return w ? static_cast<void*>(w->stream) : nullptr; }

// ---- Taumoeba (src/sim/predation.cu) ---------------------------------------------
int32_t astro_taumoeba_spawn(astro_world* w, int32_t count, uint64_t seed) {
    if (!w || count < 0) return ASTRO_INVALID_ARGUMENT;
    if (count == 0) return ASTRO_OK;
    TauCtr tc{};
    int32_t st = read_tau(w, tc);
    if (st != ASTRO_OK) return st;
    if (static_cast<int64_t>(tc.count) + count > tc.capacity)
        return fail(w, ASTRO_CAPACITY_EXCEEDED, "taumoeba spawn exceeds capacity");
    const astro_world_desc& d = w->desc;
    launch_tau_spawn(Launch{w->stream, nullptr}, w->tau, count, tc.count, tc.next_id, seed, d.chamber_w,
                     d.chamber_h, d.chamber_d);
    if ((st = hip_check(w, hipGetLastError(), "taumoeba spawn")) != ASTRO_OK) return st;
    tc.count += count;
    tc.next_id += static_cast<uint64_t>(count);
    if ((st = write_tau(w, tc)) != ASTRO_OK) return st;
    if (!w->tau_on) drop_graph(w);   // the captured tick has no predation stage
    w->tau_on = true;
    return ASTRO_OK;
}

int32_t astro_taumoeba_upload(astro_world* w, const astro_tau_soa* src, int32_t n) {
    if (!w || !src || n < 0) return ASTRO_INVALID_ARGUMENT;
    if (n > w->tau_cap) return fail(w, ASTRO_CAPACITY_EXCEEDED, "taumoeba upload exceeds capacity");
    int32_t st = hip_check(w, hipStreamSynchronize(w->stream), "synchronize");
    if (st != ASTRO_OK) return st;
    if (n > 0 && (st = copy_tau(w, src, n, true)) != ASTRO_OK) return st;
    TauCtr tc{};
    tc.count = n;
    tc.capacity = w->tau_cap;
    uint64_t mx = 0;
    for (int32_t i = 0; i < n; ++i) mx = src->id[i] > mx ? src->id[i] : mx;
    tc.next_id = mx + 1;
    if ((st = write_tau(w, tc)) != ASTRO_OK) return st;
    if (n > 0 && !w->tau_on) {
        drop_graph(w);
        w->tau_on = true;
    }
    return ASTRO_OK;
}

int32_t astro_taumoeba_download(astro_world* w, astro_tau_soa* dst, int32_t n, int32_t* count_out) {
    if (!w || !dst || n < 0) return ASTRO_INVALID_ARGUMENT;
    TauCtr tc{};
    int32_t st = read_tau(w, tc);
    if (st != ASTRO_OK) return st;
    if (count_out) *count_out = tc.count;
    const int32_t m = n < tc.count ? n : tc.count;
    return m > 0 ? copy_tau(w, dst, m, false) : ASTRO_OK;
}

int32_t astro_taumoeba_counters(astro_world* w, int32_t* count, int32_t* capacity, uint64_t* next_id) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    TauCtr tc{};
    const int32_t st = read_tau(w, tc);
    if (st != ASTRO_OK) return st;
    if (count) *count = tc.count;
    if (capacity) *capacity = tc.capacity;
    if (next_id) *next_id = tc.next_id;
    return ASTRO_OK;
}

int32_t astro_stage_predation(astro_world* w, double dt) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    int32_t st = ensure_hash(w);
    if (st != ASTRO_OK) return st;
    const Launch L = launch_of(w);
    if (w->fp32) launch_predation<float>(L, make_view<float>(w), w->tau, tau_tick(w, dt));
    else launch_predation<double>(L, make_view<double>(w), w->tau, tau_tick(w, dt));
    return hip_check(w, hipGetLastError(), "predation");
}

// ---- snapshots (src/sim/snapshot.cu:131-266) ------------------------------------
int32_t astro_snapshot_to_bytes(astro_world* w, const char* scenario_id, void* buf, size_t cap, size_t* n_out) {
    if (!w) return ASTRO_INVALID_ARGUMENT;
    std::vector<char> body;
    int32_t oc = 0, count = 0;
    int32_t st = serialize_body(w, body, oc, count);
    if (st != ASTRO_OK) return st;
    const size_t total = sizeof(SnapHeader) + body.size();
    if (n_out) *n_out = total;
    if (!buf) return ASTRO_OK;
    if (cap < total) return fail(w, ASTRO_CAPACITY_EXCEEDED, "snapshot buffer too small");
    SnapHeader h;
    std::memset(&h, 0, sizeof(h));
    h.magic = SNAP_MAGIC;
    h.version = SNAP_VERSION;
    h.seed = w->desc.seed;
    h.tick = w->tick;
    h.sim_time_s = w->sim_time;
    h.cell_count = count;
    h.cell_capacity = w->capacity;
    TauCtr tc{};
    if ((st = read_tau(w, tc)) != ASTRO_OK) return st;
    h.taumoeba_count = tc.count;
    h.taumoeba_capacity = w->tau_cap;
    h.n_temp = snap_grid_n(w->f[ASTRO_FIELD_TEMPERATURE]);
    h.n_co2 = snap_grid_n(w->f[ASTRO_FIELD_CO2]);
    h.n_n2 = snap_grid_n(w->f[ASTRO_FIELD_N2]);
    h.chamber_w = w->desc.chamber_w;
    h.chamber_h = w->desc.chamber_h;
    h.chamber_d = w->desc.chamber_d;
    h.physics_rate = w->physics_rate;
    h.biology_rate = w->biology_rate;
    Counters c{};
    if ((st = read_counters(w, c)) != ASTRO_OK) return st;
    h.next_cell_id = c.next_id;
    h.override_count = oc;
    h.non_canon_run = (oc > 0 || w->non_canon) ? 1 : 0;
    if (scenario_id) std::strncpy(h.scenario_id, scenario_id, sizeof(h.scenario_id) - 1);
    std::strncpy(h.build_describe, "astro_hot", sizeof(h.build_describe) - 1);
    h.state_hash = fnv1a64(body.data(), body.size(), FNV_OFFSET);
    std::memcpy(buf, &h, sizeof(h));
    if (!body.empty()) std::memcpy(static_cast<char*>(buf) + sizeof(h), body.data(), body.size());
    return ASTRO_OK;
}

int32_t astro_snapshot_save(astro_world* w, const char* path, const char* scenario_id) {
    if (!w || !path) return ASTRO_INVALID_ARGUMENT;
    size_t n = 0;
    int32_t st = astro_snapshot_to_bytes(w, scenario_id, nullptr, 0, &n);
    if (st != ASTRO_OK) return st;
    std::vector<char> bytes(n);
    if ((st = astro_snapshot_to_bytes(w, scenario_id, bytes.data(), n, &n)) != ASTRO_OK) return st;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return fail(w, ASTRO_FILE_IO_ERROR, "snapshot: cannot open file for write");
    const bool wrote = std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
    const bool closed = std::fclose(f) == 0;
    return wrote && closed ? ASTRO_OK : fail(w, ASTRO_FILE_IO_ERROR, "snapshot: write failed");
}

int32_t astro_snapshot_from_bytes(const astro_world_desc* base, const void* data, size_t n, astro_world** out) {
    if (!data || !out) return ASTRO_INVALID_ARGUMENT;
    *out = nullptr;
    if (n < sizeof(SnapHeader)) return ASTRO_FILE_IO_ERROR;
    SnapHeader h;
    std::memcpy(&h, data, sizeof(h));
    if (h.magic != SNAP_MAGIC) return ASTRO_INVALID_ARGUMENT;
    if (h.version != SNAP_VERSION) return ASTRO_CONTRACT_VERSION_MISMATCH;
    const char* body = static_cast<const char*>(data) + sizeof(h);
    const size_t body_len = n - sizeof(h);
    if (fnv1a64(body, body_len, FNV_OFFSET) != h.state_hash) return ASTRO_INVALID_ARGUMENT;   // corrupt
    if (h.cell_count < 0 || h.cell_count > h.cell_capacity || h.override_count < 0) return ASTRO_INVALID_ARGUMENT;
    if (h.taumoeba_count < 0 || h.taumoeba_count > h.taumoeba_capacity) return ASTRO_INVALID_ARGUMENT;
    const char* cur = body;
    const char* end = body + body_len;
    // The description: the caller's (motion configuration and grids are not
    // serialized), with the header's chamber, capacity, seed, clock rates and
    // overrides.
    astro_world_desc d;
    if (base) d = *base;
    else astro_world_desc_default(&d);
    d.capacity = h.cell_capacity;
    d.tau_capacity = h.taumoeba_capacity;
    d.chamber_w = h.chamber_w;
    d.chamber_h = h.chamber_h;
    d.chamber_d = h.chamber_d;
    d.seed = h.seed;
    d.physics_rate = h.physics_rate;
    d.biology_rate = h.biology_rate;
    for (int32_t i = 0; i < h.override_count; ++i) {
        if (static_cast<size_t>(end - cur) < sizeof(SnapOverride)) return ASTRO_INVALID_ARGUMENT;
        SnapOverride o;
        std::memcpy(&o, cur, sizeof(o));
        cur += sizeof(o);
        if (std::strncmp(o.key, "PETROVA_MAX_POWER", sizeof(o.key)) == 0) d.petrova_max_power = o.value;
        else if (std::strncmp(o.key, "PETROVA_FLASH_POWER", sizeof(o.key)) == 0) d.petrova_flash_power = o.value;
        else if (std::strncmp(o.key, "CO2_MASS_PER_DIVISION", sizeof(o.key)) == 0) d.co2_mass_per_division = o.value;
    }
    astro_world* w = nullptr;
    int32_t st = astro_world_create(&d, &w);
    if (st != ASTRO_OK) return st;
    // a malformed body leaves no world behind
    auto bail = [&](int32_t code) {
        astro_world_destroy(w);
        return code;
    };
    const int32_t want[3] = {h.n_temp, h.n_co2, h.n_n2};
    const int fields[3] = {ASTRO_FIELD_TEMPERATURE, ASTRO_FIELD_CO2, ASTRO_FIELD_N2};
    for (int k = 0; k < 3; ++k)
        if (snap_grid_n(w->f[fields[k]]) != want[k]) return bail(ASTRO_INVALID_ARGUMENT);
    const int32_t count = h.cell_count;
    if (count > 0) {
        HostSoa soa(count);
        bool ok = true;
        soa.each([&](auto& a) {
            const size_t bytes = sizeof(a[0]) * a.size();
            if (!ok || static_cast<size_t>(end - cur) < bytes) {
                ok = false;
                return;
            }
            std::memcpy(a.data(), cur, bytes);
            cur += bytes;
        });
        if (!ok) return bail(ASTRO_INVALID_ARGUMENT);
        astro_cell_soa v = soa.view();
        if ((st = astro_cells_upload(w, &v, count)) != ASTRO_OK) return bail(st);
    }
    if (h.taumoeba_count > 0) {
        HostTau tau(static_cast<size_t>(h.taumoeba_count));
        bool ok = true;
        tau.each([&](auto& a) {
            const size_t bytes = sizeof(a[0]) * a.size();
            if (!ok || static_cast<size_t>(end - cur) < bytes) {
                ok = false;
                return;
            }
            std::memcpy(a.data(), cur, bytes);
            cur += bytes;
        });
        if (!ok) return bail(ASTRO_INVALID_ARGUMENT);
        astro_tau_soa v = tau.view();
        if ((st = astro_taumoeba_upload(w, &v, h.taumoeba_count)) != ASTRO_OK) return bail(st);
    }
    for (int k = 0; k < 3; ++k) {
        const FieldHost& g = w->f[fields[k]];
        if (!g.on) continue;
        const size_t bytes = sizeof(float) * g.count();
        if (static_cast<size_t>(end - cur) < bytes) return bail(ASTRO_INVALID_ARGUMENT);
        std::vector<float> v(g.count());
        std::memcpy(v.data(), cur, bytes);
        cur += bytes;
        if ((st = astro_field_upload(w, fields[k], v.data(), v.size())) != ASTRO_OK) return bail(st);
    }
    // bytes past the fields are left unread, as by the reference's loader
    astro_counters c{};
    if ((st = astro_world_get_counters(w, &c)) != ASTRO_OK) return bail(st);
    c.next_id = h.next_cell_id;
    c.tick = h.tick;
    c.sim_time_s = h.sim_time_s;
    if ((st = astro_world_set_counters(w, &c)) != ASTRO_OK) return bail(st);
    w->non_canon = h.non_canon_run != 0;
    *out = w;
    return ASTRO_OK;
}

int32_t astro_snapshot_load(const astro_world_desc* base, const char* path, astro_world** out) {
    if (!path || !out) return ASTRO_INVALID_ARGUMENT;
    *out = nullptr;
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return ASTRO_FILE_IO_ERROR;
    std::vector<char> all;
    char buf[65536];
    size_t r = 0;
    while ((r = std::fread(buf, 1, sizeof(buf), f)) > 0) all.insert(all.end(), buf, buf + r);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) return ASTRO_FILE_IO_ERROR;
    return astro_snapshot_from_bytes(base, all.data(), all.size(), out);
}

// ---- ensembles ----------------------------------------------------------------
int32_t astro_ensemble_create(astro_world* const* chambers, int32_t n, astro_ensemble** out) {
    if (!chambers || n <= 0 || !out) return ASTRO_INVALID_ARGUMENT;
    *out = nullptr;
    for (int32_t i = 0; i < n; ++i) {
        if (!chambers[i]) return ASTRO_INVALID_ARGUMENT;
        for (int32_t j = 0; j < i; ++j)
            if (chambers[j] == chambers[i]) return ASTRO_INVALID_ARGUMENT;
        if (chambers[i]->device != chambers[0]->device) return ASTRO_INVALID_ARGUMENT;
    }
    auto e = std::make_unique<astro_ensemble>();
    e->worlds.assign(chambers, chambers + n);
    int32_t st = check_shape(e.get());
    if (st != ASTRO_OK) return st;
    if ((st = echeck(e.get(), hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "stream")) != ASTRO_OK)
        return st;
    if ((st = echeck(e.get(), hipEventCreateWithFlags(&e->done, hipEventDisableTiming), "event")) != ASTRO_OK) {
        astro_ensemble_destroy(e.release());
        return st;
    }
    e->ready.assign(n, nullptr);
    for (int32_t i = 0; i < n; ++i)
        if ((st = echeck(e.get(), hipEventCreateWithFlags(&e->ready[i], hipEventDisableTiming), "event")) != ASTRO_OK) {
            astro_ensemble_destroy(e.release());
            return st;
        }
    e->prof = std::make_unique<EventProfiler>();
    e->prof->stream = e->stream;
    *out = e.release();
    return ASTRO_OK;
}

void astro_ensemble_destroy(astro_ensemble* e) {
    if (!e) return;
    if (e->stream) hipStreamSynchronize(e->stream);
    drop_ensemble_graph(e);
    if (e->eager) hipFree(e->eager);
    e->prof.reset();
    for (hipEvent_t ev : e->ready)
        if (ev) hipEventDestroy(ev);
    if (e->done) hipEventDestroy(e->done);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
}

const char* astro_ensemble_last_error(const astro_ensemble* e) { return e ? e->err.c_str() : ""; }

int32_t astro_ensemble_size(const astro_ensemble* e) { return e ? static_cast<int32_t>(e->worlds.size()) : 0; }

astro_world* astro_ensemble_chamber(astro_ensemble* e, int32_t i) {
    if (!e || i < 0 || i >= static_cast<int32_t>(e->worlds.size())) return nullptr;
    return e->worlds[i];
}

int32_t astro_ensemble_step(astro_ensemble* e, int32_t n_ticks) {
    if (!e || n_ticks < 0) return ASTRO_INVALID_ARGUMENT;
    if (n_ticks == 0) return ASTRO_OK;
    int32_t st = check_shape(e);
    if (st != ASTRO_OK) return st;
    // Order the step after whatever each chamber's own stream holds.
    for (size_t i = 0; i < e->worlds.size(); ++i) {
        if ((st = echeck(e, hipEventRecord(e->ready[i], e->worlds[i]->stream), "record")) != ASTRO_OK) return st;
        if ((st = echeck(e, hipStreamWaitEvent(e->stream, e->ready[i], 0), "wait")) != ASTRO_OK) return st;
    }
    int32_t done = 0;
    if (!e->prof_on) {
        const int k = ticks_per_graph(e->worlds[0]);
        if (n_ticks >= k) {
            if (!e->gexec || e->graph_ticks != k || chambers_changed(e)) {
                if ((st = capture_step(e)) != ASTRO_OK) return st;
            }
            while (n_ticks - done >= e->graph_ticks) {
                if ((st = echeck(e, hipGraphLaunch(e->gexec, e->stream), "graph launch")) != ASTRO_OK) return st;
                for (astro_world* w : e->worlds) advance_clock(w, e->graph_ticks);
                done += e->graph_ticks;
            }
            for (astro_world* w : e->worlds) w->hash_valid = false;
        }
    }
    while (done < n_ticks) {
        if ((st = eager_tick(e)) != ASTRO_OK) return st;
        for (astro_world* w : e->worlds) advance_clock(w, 1);
        ++done;
        // A lone tick leaves a field's ping-pong swapped when its substep count
        // is odd; graphs captured at the other parity no longer apply.
        if (ticks_per_graph(e->worlds[0]) == 2) {
            drop_ensemble_graph(e);
            for (astro_world* w : e->worlds) drop_graph(w);
        }
    }
    // Each chamber's own stream now waits for the step.
    if ((st = echeck(e, hipEventRecord(e->done, e->stream), "record")) != ASTRO_OK) return st;
    for (astro_world* w : e->worlds)
        if ((st = echeck(e, hipStreamWaitEvent(w->stream, e->done, 0), "wait")) != ASTRO_OK) return st;
    return ASTRO_OK;
}

int32_t astro_ensemble_sync(astro_ensemble* e) {
    if (!e) return ASTRO_INVALID_ARGUMENT;
    int32_t st = echeck(e, hipStreamSynchronize(e->stream), "synchronize");
    if (st != ASTRO_OK) return st;
    for (astro_world* w : e->worlds)
        if ((st = check_device_errors(w)) != ASTRO_OK) return efail(e, st, w->err);
    return ASTRO_OK;
}

// Launch shapes tried by the autotuner: total blocks of a grid-stride stage
// (0: one block per tile of each chamber's capacity) and lanes per x row of
// the float4 FTCS (0: the widest row group nx fills).
static const int TUNE_BLOCKS[] = {0, 1024, 2048, 4096, 8192, 16384};
static const int TUNE_LANES[] = {0, 8, 16, 32, 64};

int32_t astro_ensemble_autotune(astro_ensemble* e, int32_t ticks) {
    if (!e || ticks <= 0) return ASTRO_INVALID_ARGUMENT;
    int32_t st = check_shape(e);
    if (st != ASTRO_OK) return st;
    const bool was_on = e->prof_on;
    // One trial: `ticks` eager ticks with per-stage event timing; returns
    // milliseconds per stage.
    const auto trial = [&](std::map<std::string, double>& ms) -> int32_t {
        e->prof->drain();
        e->prof->totals.clear();
        e->prof->order.clear();
        e->prof_on = true;
        const int32_t rc = astro_ensemble_step(e, ticks);
        e->prof_on = was_on;
        if (rc != ASTRO_OK) return rc;
        e->prof->drain();
        ms.clear();
        for (const auto& [name, v] : e->prof->totals) ms[name] = v.second;
        return ASTRO_OK;
    };
    std::map<std::string, std::pair<double, int>> best_blocks, best_lanes;
    e->stage_blocks.clear();
    e->ftcs_lanes.clear();
    std::map<std::string, double> ms;
    for (const int b : TUNE_BLOCKS) {
        e->stride_blocks = b;
        if ((st = trial(ms)) != ASTRO_OK) return st;
        for (const auto& [name, t] : ms) {
            auto it = best_blocks.find(name);
            if (it == best_blocks.end() || t < it->second.first) best_blocks[name] = {t, b};
        }
    }
    e->stride_blocks = 2048;
    for (const auto& [name, v] : best_blocks) e->stage_blocks[name] = v.second;
    for (const int l : TUNE_LANES) {
        for (const char* f : {"ftcs_T", "ftcs_CO2", "ftcs_N2"}) e->ftcs_lanes[f] = l;
        if ((st = trial(ms)) != ASTRO_OK) return st;
        for (const auto& [name, t] : ms) {
            if (name.rfind("ftcs_", 0) != 0) continue;
            auto it = best_lanes.find(name);
            if (it == best_lanes.end() || t < it->second.first) best_lanes[name] = {t, l};
        }
    }
    e->ftcs_lanes.clear();
    for (const auto& [name, v] : best_lanes) e->ftcs_lanes[name] = v.second;
    drop_ensemble_graph(e);
    e->prof->totals.clear();
    e->prof->order.clear();
    return ASTRO_OK;
}

int32_t astro_ensemble_tuning(const astro_ensemble* e, char* out, int32_t len) {
    if (!e || !out || len <= 0) return ASTRO_INVALID_ARGUMENT;
    std::string s = "{\"default_blocks\": " + std::to_string(e->stride_blocks) + ", \"stage_blocks\": {";
    bool first = true;
    for (const auto& [name, b] : e->stage_blocks) {
        s += (first ? "\"" : ", \"") + name + "\": " + std::to_string(b);
        first = false;
    }
    s += "}, \"ftcs_lanes\": {";
    first = true;
    for (const auto& [name, l] : e->ftcs_lanes) {
        s += (first ? "\"" : ", \"") + name + "\": " + std::to_string(l);
        first = false;
    }
    s += "}}";
    if (static_cast<int32_t>(s.size()) >= len) return ASTRO_CAPACITY_EXCEEDED;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return ASTRO_OK;
}

int32_t astro_ensemble_profile_enable(astro_ensemble* e, int32_t on) {
    if (!e) return ASTRO_INVALID_ARGUMENT;
    e->prof_on = on != 0;
    return ASTRO_OK;
}

int32_t astro_ensemble_profile_read(astro_ensemble* e, char* names, int32_t stride, int64_t* launches,
                                    double* total_ms, int32_t max, int32_t* n_out) {
    if (!e) return ASTRO_INVALID_ARGUMENT;
    e->prof->drain();
    int32_t k = 0;
    for (const std::string& name : e->prof->order) {
        if (k >= max) break;
        const auto& v = e->prof->totals[name];
        if (names && stride > 0) {
            std::strncpy(names + static_cast<size_t>(k) * stride, name.c_str(), stride - 1);
            names[static_cast<size_t>(k) * stride + stride - 1] = '\0';
        }
        if (launches) launches[k] = v.first;
        if (total_ms) total_ms[k] = v.second;
        ++k;
    }
    if (n_out) *n_out = k;
    return ASTRO_OK;
}

int32_t astro_ensemble_profile_reset(astro_ensemble* e) {
    if (!e) return ASTRO_INVALID_ARGUMENT;
    e->prof->drain();
    e->prof->totals.clear();
    e->prof->order.clear();
    return ASTRO_OK;
}

void* astro_ensemble_stream(astro_ensemble* e) { return e ? static_cast<void*>(e->stream) : nullptr; }

}  // extern "C"

// ---------------------------------------------------------------------------
// Counter-based streams, evaluated on the device for known-answer checks.
// ---------------------------------------------------------------------------
namespace {

__global__ void k_philox_blocks(const uint32_t* words, int32_t n, uint32_t* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* w = words + 6 * static_cast<size_t>(i);
    const uint4 r = philox4x32_10(uint4{w[0], w[1], w[2], w[3]}, w[4], w[5]);
    uint32_t* o = out + 4 * static_cast<size_t>(i);
    o[0] = r.x; o[1] = r.y; o[2] = r.z; o[3] = r.w;
}

__global__ void k_philox_stream(uint64_t seed, uint64_t id, uint64_t tick, uint32_t stage, int32_t n,
                                uint32_t* out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    Philox r(seed, stage, tick, id);
    for (int32_t k = 0; k < n; ++k) out[k] = r.next();
}

// Runs `launch` on a device copy of `in` (in_words uint32) writing out_words
// uint32 back to `out`, on the current device's null stream.
template <class F>
int32_t run_words(const uint32_t* in, size_t in_words, uint32_t* out, size_t out_words, F launch) {
    uint32_t *din = nullptr, *dout = nullptr;
    int32_t st = ASTRO_OK;
    if (in_words && hipMalloc(&din, in_words * 4) != hipSuccess) return ASTRO_OUT_OF_MEMORY;
    if (hipMalloc(&dout, std::max<size_t>(out_words, 1) * 4) != hipSuccess) st = ASTRO_OUT_OF_MEMORY;
    if (st == ASTRO_OK && in_words && hipMemcpy(din, in, in_words * 4, hipMemcpyHostToDevice) != hipSuccess)
        st = ASTRO_DEVICE_ERROR;
    if (st == ASTRO_OK) {
        launch(din, dout);
        if (hipGetLastError() != hipSuccess) st = ASTRO_DEVICE_ERROR;
    }
    if (st == ASTRO_OK && hipMemcpy(out, dout, out_words * 4, hipMemcpyDeviceToHost) != hipSuccess)
        st = ASTRO_DEVICE_ERROR;
    if (din) hipFree(din);
    if (dout) hipFree(dout);
    return st;
}

}  // namespace

extern "C" {

int32_t astro_philox_blocks(const uint32_t* words, int32_t n, uint32_t* out) {
    if (n < 0 || (n > 0 && (!words || !out))) return ASTRO_INVALID_ARGUMENT;
    if (n == 0) return ASTRO_OK;
    const size_t m = static_cast<size_t>(n);
    return run_words(words, 6 * m, out, 4 * m, [&](const uint32_t* din, uint32_t* dout) {
        k_philox_blocks<<<(n + BLOCK - 1) / BLOCK, BLOCK>>>(din, n, dout);
    });
}

int32_t astro_philox_stream(uint64_t seed, uint64_t id, uint64_t tick, int32_t stage, int32_t n, uint32_t* out) {
    if (n < 0 || stage < 0 || stage > 0xFFFF || (n > 0 && !out)) return ASTRO_INVALID_ARGUMENT;
    if (n == 0) return ASTRO_OK;
    return run_words(nullptr, 0, out, static_cast<size_t>(n), [&](const uint32_t*, uint32_t* dout) {
        k_philox_stream<<<1, 1>>>(seed, id, tick, static_cast<uint32_t>(stage), n, dout);
    });
}

}  // extern "C"

