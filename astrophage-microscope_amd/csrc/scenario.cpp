// scenario.cpp -- scenarios over the MI355X world (include/astro_scenario.h).
//
// Host code only: parsing, instantiation, the per-tick driving script and the
// acceptance runner all go through the world's public C ABI, exactly as a
// reference caller would. Semantics follow the reference's scenario layer
// (src/sim/scenario.cpp), acceptance evaluation (src/sim/accept.cpp) and
// headless runner (tools/headless.cpp:157-250); each function cites the part
// it follows.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <utility>
#include <vector>

#include "astro_canon.h"
#include "astro_scenario.h"
#include "jsonc.h"

namespace {

// Truncating, always-terminated copy into a fixed char field.
void copy_text(char* dst, size_t cap, const std::string& v) {
    const size_t n = std::min(v.size(), cap - 1);
    if (n) std::memcpy(dst, v.data(), n);
    dst[n] = '\0';
}

// A name -> enum table; unknown names give the fallback (or fail when strict).
struct Named {
    const char* name;
    int32_t value;
};
template <size_t N>
int32_t pick(const Named (&table)[N], const std::string& s, int32_t fallback, bool* found = nullptr) {
    for (const Named& e : table)
        if (s == e.name) {
            if (found) *found = true;
            return e.value;
        }
    if (found) *found = false;
    return fallback;
}

// The string vocabularies of src/sim/scenario.cpp:28-137.
const Named BOUNDARIES[] = {{"reflecting", ASTRO_BOUNDARY_REFLECTING},
                            {"periodic", ASTRO_BOUNDARY_PERIODIC},
                            {"absorbing", ASTRO_BOUNDARY_ABSORBING}};
const Named THERMAL_BCS[] = {{"dirichlet", ASTRO_BC_DIRICHLET}, {"neumann", ASTRO_BC_NEUMANN},
                             {"robin", ASTRO_BC_ROBIN}};
const Named DENSITIES[] = {{"canon-mass", ASTRO_DENSITY_CANON_MASS}, {"water-density", ASTRO_DENSITY_WATER}};
const Named DISPOSITIONS[] = {{"void", ASTRO_STORE_VOID}, {"flash", ASTRO_STORE_FLASH},
                              {"retain", ASTRO_STORE_RETAIN}};
const Named CLOCKS[] = {{"realtime", ASTRO_CLOCK_REALTIME}, {"motion", ASTRO_CLOCK_MOTION},
                        {"metabolic", ASTRO_CLOCK_METABOLIC}, {"generational", ASTRO_CLOCK_GENERATIONAL},
                        {"custom", ASTRO_CLOCK_CUSTOM}};
const Named PLACEMENTS[] = {{"uniform", ASTRO_PLACE_UNIFORM}, {"gaussian", ASTRO_PLACE_GAUSSIAN},
                            {"grid", ASTRO_PLACE_GRID}, {"disc", ASTRO_PLACE_DISC}};
const Named DISTRIBUTIONS[] = {{"constant", ASTRO_DIST_CONSTANT}, {"uniform", ASTRO_DIST_UNIFORM},
                               {"normal", ASTRO_DIST_NORMAL}};
const Named ORGANISMS[] = {{"astrophage", ASTRO_ORGANISM_ASTROPHAGE}, {"taumoeba", ASTRO_ORGANISM_TAUMOEBA}};
// ToolBit (contracts/scenario_v3.h:39-43)
const Named TOOLS[] = {{"heat", 1 << 0},       {"chill", 1 << 1},      {"illuminate", 1 << 2},
                       {"inject_co2", 1 << 3}, {"inject_n2", 1 << 4},  {"seed_cells", 1 << 5},
                       {"seed_taumoeba", 1 << 6}, {"kill", 1 << 7}, {"charge_beam", 1 << 8}};
const Named STIMULI[] = {{"heat", ASTRO_STIM_HEAT},           {"chill", ASTRO_STIM_CHILL},
                         {"inject_co2", ASTRO_STIM_INJECT_CO2}, {"inject_n2", ASTRO_STIM_INJECT_N2},
                         {"set_light", ASTRO_STIM_SET_LIGHT}, {"set_n2", ASTRO_STIM_SET_N2},
                         {"seed_cells", ASTRO_STIM_SEED_CELLS}, {"flash", ASTRO_STIM_FLASH}};
const Named OPS[] = {{"==", ASTRO_OP_EQ}, {"!=", ASTRO_OP_NE}, {"<", ASTRO_OP_LT}, {"<=", ASTRO_OP_LE},
                     {">", ASTRO_OP_GT},  {">=", ASTRO_OP_GE}, {"~=", ASTRO_OP_APPROX}};
const Named METRICS[] = {{"awake_fraction", ASTRO_METRIC_AWAKE_FRACTION},
                         {"medium_temp_mean", ASTRO_METRIC_MEDIUM_TEMP_MEAN},
                         {"medium_temp_max", ASTRO_METRIC_MEDIUM_TEMP_MAX},
                         {"boil_event_count", ASTRO_METRIC_BOIL_EVENT_COUNT},
                         {"population_live", ASTRO_METRIC_POPULATION_LIVE},
                         {"population_dead", ASTRO_METRIC_POPULATION_DEAD},
                         {"mean_charge", ASTRO_METRIC_MEAN_CHARGE},
                         {"total_energy_j", ASTRO_METRIC_TOTAL_ENERGY_J},
                         {"mean_tau_tolerance", ASTRO_METRIC_MEAN_TAU_TOLERANCE},
                         {"max_tau_tolerance", ASTRO_METRIC_MAX_TAU_TOLERANCE},
                         {"charge_depth_correlation", ASTRO_METRIC_CHARGE_DEPTH_CORRELATION},
                         {"charge_height_correlation", ASTRO_METRIC_CHARGE_HEIGHT_CORRELATION},
                         {"rise_velocity_empty", ASTRO_METRIC_RISE_VELOCITY_EMPTY},
                         {"fall_velocity_full", ASTRO_METRIC_FALL_VELOCITY_FULL},
                         {"doubling_time_s", ASTRO_METRIC_DOUBLING_TIME_S},
                         {"impulse_per_cycle", ASTRO_METRIC_IMPULSE_PER_CYCLE}};

// scenario_parse's field reads (src/sim/scenario.cpp:139-318), onto a
// zeroed scenario.
int32_t read_scenario(const jsonc::Node& root, astro_scenario& o) {
    o.schema = static_cast<int32_t>(root.num("schema", 1));
    copy_text(o.id, sizeof(o.id), root.str("id", ""));
    copy_text(o.title, sizeof(o.title), root.str("title", ""));
    copy_text(o.blurb, sizeof(o.blurb), root.str("blurb", ""));
    o.seed = static_cast<uint64_t>(root.num("seed", 20260802.0));

    o.chamber_w = ASTRO_CHAMBER_W;
    o.chamber_h = ASTRO_CHAMBER_H;
    o.chamber_d = ASTRO_CHAMBER_D;
    if (const jsonc::Node* ch = root.at("chamber")) {
        o.chamber_w = ch->num("w", ASTRO_CHAMBER_W);
        o.chamber_h = ch->num("h", ASTRO_CHAMBER_H);
        o.chamber_d = ch->num("d", ASTRO_CHAMBER_D);
        o.boundary_x = pick(BOUNDARIES, ch->str("boundary_x", "reflecting"), ASTRO_BOUNDARY_REFLECTING);
        o.boundary_y = pick(BOUNDARIES, ch->str("boundary_y", "reflecting"), ASTRO_BOUNDARY_REFLECTING);
    }

    o.temp_init = ASTRO_AMBIENT_TEMP_DEFAULT;
    o.ambient_temp = ASTRO_AMBIENT_TEMP_DEFAULT;
    o.thermal_bc = ASTRO_BC_ROBIN;
    o.density_model = ASTRO_DENSITY_CANON_MASS;
    o.store_disposition = ASTRO_STORE_VOID;
    o.thermal_active = 1;
    if (const jsonc::Node* md = root.at("medium")) {
        o.temp_init = md->num("temp_init", ASTRO_AMBIENT_TEMP_DEFAULT);
        o.co2_init = md->num("co2_init", 0.0);
        o.n2_init = md->num("n2_init", 0.0);
        o.ambient_temp = md->num("ambient_temp", ASTRO_AMBIENT_TEMP_DEFAULT);
        o.thermal_bc = pick(THERMAL_BCS, md->str("thermal_bc", "robin"), ASTRO_BC_ROBIN);
        o.density_model = pick(DENSITIES, md->str("density_model", "canon-mass"), ASTRO_DENSITY_CANON_MASS);
        o.store_disposition = pick(DISPOSITIONS, md->str("store_disposition", "void"), ASTRO_STORE_VOID);
        o.gravity_axis = md->str("gravity_axis", "y") == "z" ? ASTRO_GRAVITY_Z : ASTRO_GRAVITY_Y;
        o.thermal_active = md->flag("thermal_active", true) ? 1 : 0;
    }

    if (const jsonc::Node* lt = root.at("light")) {
        o.ambient_irradiance = static_cast<float>(lt->num("ambient", 0.0));
        int n = 0;
        for (const jsonc::Node& src : lt->list("sources")) {
            if (n >= ASTRO_SCENARIO_MAX_LIGHTS) break;
            astro_light_source& l = o.lights[n++];
            l.dir_x = static_cast<float>(src.num("dir_x", 1.0));
            l.dir_y = static_cast<float>(src.num("dir_y", 0.0));
            l.irradiance = static_cast<float>(src.num("irradiance", 0.0));
            l.wavelength = static_cast<float>(src.num("wavelength", 0.0));
            l.enabled = 1;
        }
    }

    int np = 0;
    for (const jsonc::Node& p : root.list("populations")) {
        if (np >= ASTRO_SCENARIO_MAX_POPULATIONS) break;
        astro_population_spec& ps = o.populations[np++];
        ps.kind = pick(ORGANISMS, p.str("kind", "astrophage"), ASTRO_ORGANISM_ASTROPHAGE);
        ps.count = static_cast<int32_t>(p.num("count", 0));
        ps.placement = pick(PLACEMENTS, p.str("placement", "uniform"), ASTRO_PLACE_UNIFORM);
        ps.place_x = p.num("place_x", 0.0);
        ps.place_y = p.num("place_y", 0.0);
        ps.place_radius = p.num("place_radius", 0.0);
        ps.charge_dist = ASTRO_DIST_CONSTANT;
        if (const jsonc::Node* cg = p.at("charge")) {
            ps.charge_dist = pick(DISTRIBUTIONS, cg->str("dist", "constant"), ASTRO_DIST_CONSTANT);
            ps.charge_a = cg->num("value", cg->num("a", 0.0));
            ps.charge_b = cg->num("b", 0.0);
        }
        ps.awake = p.flag("awake", false) ? 1 : 0;
    }
    o.population_count = np;

    o.compaction = root.flag("compaction", false) ? 1 : 0;
    o.tau_compaction = root.flag("tau_compaction", false) ? 1 : 0;

    o.clock = ASTRO_CLOCK_REALTIME;
    o.physics_rate = 1.0;
    o.biology_rate = 1.0;
    if (const jsonc::Node* ck = root.at("clock")) {
        o.clock = pick(CLOCKS, ck->str("preset", "realtime"), ASTRO_CLOCK_REALTIME);
        o.physics_rate = ck->num("physics_rate", 1.0);
        o.biology_rate = ck->num("biology_rate", 1.0);
    }

    for (const jsonc::Node& t : root.list("tools"))
        if (t.kind == jsonc::Node::STR) o.tools |= static_cast<uint32_t>(pick(TOOLS, t.s, 0));

    int nd = 0;
    for (const jsonc::Node& e : root.list("drive")) {
        if (nd >= ASTRO_SCENARIO_MAX_STIMULI) break;
        bool known = false;
        const int32_t kind = pick(STIMULI, e.str("kind", ""), ASTRO_STIM_HEAT, &known);
        if (!known) return ASTRO_INVALID_ARGUMENT;   // a dropped stimulus would change the run
        astro_stimulus& st = o.drive[nd++];
        st.kind = kind;
        st.t0_s = e.num("t0_s", 0.0);
        st.t1_s = e.num("t1_s", 0.0);
        st.x = e.num("x", 0.0);
        st.y = e.num("y", 0.0);
        st.radius = e.num("radius", 0.0);
        st.v0 = e.num("v0", e.num("value", 0.0));
        st.v1 = e.num("v1", e.num("value", st.v0));
    }
    o.drive_count = nd;

    // Overrides are carried, not applied, as in the reference.
    if (const jsonc::Node* po = root.at("param_overrides")) {
        int n = 0;
        if (po->kind == jsonc::Node::OBJ)
            for (const auto& m : po->members) {
                if (n >= ASTRO_SCENARIO_MAX_OVERRIDES) break;
                copy_text(o.param_overrides[n].key, sizeof(o.param_overrides[n].key), m.first);
                o.param_overrides[n].value = m.second.as_num(0.0);
                ++n;
            }
        o.override_count = n;
    }

    if (const jsonc::Node* ob = root.at("objective")) {
        copy_text(o.objective_text, sizeof(o.objective_text), ob->str("text", ""));
        o.run_duration_s = ob->num("duration_s", 0.0);
        int na = 0;
        for (const jsonc::Node& c : ob->list("accept")) {
            if (na >= ASTRO_SCENARIO_MAX_ACCEPT) break;
            bool known = false;
            const int32_t metric = pick(METRICS, c.str("metric", ""), ASTRO_METRIC_COUNT, &known);
            if (!known) return ASTRO_INVALID_ARGUMENT;
            astro_accept_check& k = o.accept[na++];
            k.metric = metric;
            k.op = pick(OPS, c.str("op", "=="), ASTRO_OP_EQ);
            k.value = c.num("value", 0.0);
            k.tol = c.num("tol", 0.0);
            k.after_s = c.num("after_s", 0.0);
            k.tol_absolute = c.flag("tol_absolute", false) ? 1 : 0;
        }
        o.accept_count = na;
    }
    return ASTRO_OK;
}

// Is stimulus `st` applied at a tick starting at sim time t
// (src/sim/scenario.cpp:404-409)?
bool stimulus_on(const astro_stimulus& st, double t) {
    const bool open_ended = st.t1_s <= st.t0_s;
    return t >= st.t0_s && (open_ended || t < st.t1_s);
}

// ---- acceptance (src/sim/accept.cpp) -----------------------------------------

// Pearson correlation; 0 when undefined (src/sim/accept.cpp:19-34).
double correlation(const std::vector<double>& a, const std::vector<double>& b) {
    const size_t n = a.size();
    if (n < 2 || b.size() != n) return 0.0;
    double ma = 0.0, mb = 0.0;
    for (size_t i = 0; i < n; ++i) {
        ma += a[i];
        mb += b[i];
    }
    ma /= static_cast<double>(n);
    mb /= static_cast<double>(n);
    double sab = 0.0, saa = 0.0, sbb = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double da = a[i] - ma, db = b[i] - mb;
        sab += da * db;
        saa += da * da;
        sbb += db * db;
    }
    if (saa <= 0.0 || sbb <= 0.0) return 0.0;
    return sab / std::sqrt(saa * sbb);
}

// Positions and energies of slots [0, count) (the reference's
// cell_store_download_positions / _energy).
struct CellSample {
    std::vector<double> x, y, energy;
};
int32_t sample_cells(astro_world* w, CellSample& out) {
    astro_counters c{};
    int32_t st = astro_world_get_counters(w, &c);
    if (st != ASTRO_OK) return st;
    const int32_t n = c.count;
    std::vector<uint64_t> id(n), rng(n);
    std::vector<uint32_t> flags(n);
    std::vector<uint8_t> dc(n);
    std::vector<double> z(n), vx(n), vy(n), vz(n), biomass(n), co2(n);
    std::vector<float> f32[12];
    for (auto& v : f32) v.resize(n);
    out.x.assign(n, 0.0);
    out.y.assign(n, 0.0);
    out.energy.assign(n, 0.0);
    astro_cell_soa s{id.data(),     flags.data(),      dc.data(),      out.x.data(),  out.y.data(),
                     z.data(),      vx.data(),         vy.data(),      vz.data(),     out.energy.data(),
                     f32[0].data(), f32[1].data(),     f32[2].data(),  f32[3].data(), f32[4].data(),
                     biomass.data(), co2.data(),       f32[5].data(),  rng.data(),    f32[6].data(),
                     f32[7].data(), f32[8].data(),     f32[9].data(),  f32[10].data(), f32[11].data()};
    int32_t got = 0;
    return n > 0 ? astro_cells_download(w, &s, n, &got) : ASTRO_OK;
}

// RunAggregates (src/sim/accept.h:31-53).
struct Aggregates {
    int64_t boil_events = 0;
    double max_medium_k = 0.0;
    std::vector<double> t_series, n_series;
    int vel_stage = 0;
    double vel_ta = 0.0;
    std::vector<double> vel_y_a, vel_charge_a;
    double rise_velocity_empty = 0.0, fall_velocity_full = 0.0;
};

struct Needs {
    bool velocity = false, population = false;
};

// metric_needs (src/sim/accept.cpp:113-125).
Needs needs_of(const astro_scenario& s) {
    Needs n;
    for (int i = 0; i < s.accept_count; ++i) {
        const int32_t m = s.accept[i].metric;
        if (m == ASTRO_METRIC_RISE_VELOCITY_EMPTY || m == ASTRO_METRIC_FALL_VELOCITY_FULL) n.velocity = true;
        if (m == ASTRO_METRIC_DOUBLING_TIME_S) n.population = true;
    }
    return n;
}

// aggregates_sample (src/sim/accept.cpp:127-183).
int32_t sample(Aggregates& a, const astro_stats& st, astro_world* w, const Needs& needs, double settle_s,
               double interval_s, double chamber_h) {
    a.boil_events += st.boil_events;
    a.max_medium_k = std::max(a.max_medium_k, st.max_temp_medium_k);
    if (needs.population) {
        a.t_series.push_back(st.sim_time_s);
        a.n_series.push_back(static_cast<double>(st.n_live));
    }
    if (!needs.velocity || a.vel_stage >= 2) return ASTRO_OK;
    const bool take_a = a.vel_stage == 0 && st.sim_time_s >= settle_s;
    const bool take_b = a.vel_stage == 1 && st.sim_time_s >= a.vel_ta + interval_s;
    if (!take_a && !take_b) return ASTRO_OK;
    CellSample c;
    const int32_t rc = sample_cells(w, c);
    if (rc != ASTRO_OK) return rc;
    const size_t n = c.x.size();
    if (n == 0) return ASTRO_OK;
    if (take_a) {
        // first position sample: y and charge of every cell
        a.vel_y_a = c.y;
        a.vel_charge_a.resize(n);
        for (size_t i = 0; i < n; ++i) a.vel_charge_a[i] = c.energy[i] / ASTRO_CELL_ENERGY_MAX;
        a.vel_ta = st.sim_time_s;
        a.vel_stage = 1;
        return ASTRO_OK;
    }
    // second sample: settling speed = a + b * charge fitted to the drift of
    // cells away from the y walls; a is the empty, a + b the full cell's speed
    const double half_h = 0.5 * chamber_h;
    const double margin = 3.0 * ASTRO_CELL_DIAMETER;
    const double dt = st.sim_time_s - a.vel_ta;
    const size_t m = std::min(n, a.vel_y_a.size());
    std::vector<double> v, q;
    v.reserve(m);
    q.reserve(m);
    for (size_t i = 0; i < m; ++i) {
        if (std::fabs(c.y[i]) > half_h - margin) continue;
        v.push_back(-(c.y[i] - a.vel_y_a[i]) / dt);
        q.push_back(a.vel_charge_a[i]);
    }
    const size_t k = v.size();
    if (k >= 2 && dt > 0.0) {
        double mq = 0.0, mv = 0.0;
        for (size_t i = 0; i < k; ++i) {
            mq += q[i];
            mv += v[i];
        }
        mq /= static_cast<double>(k);
        mv /= static_cast<double>(k);
        double sqv = 0.0, sqq = 0.0;
        for (size_t i = 0; i < k; ++i) {
            const double dq = q[i] - mq;
            sqv += dq * (v[i] - mv);
            sqq += dq * dq;
        }
        const double b = sqq > 0.0 ? sqv / sqq : 0.0;
        const double a0 = mv - b * mq;
        a.rise_velocity_empty = a0;
        a.fall_velocity_full = a0 + b;
    }
    a.vel_stage = 2;
    return ASTRO_OK;
}

// measure_doubling_time (src/sim/accept.cpp:85-108): ln N against t below 60 %
// of the peak, reported in culture time.
double doubling_time(const Aggregates& a, double biology_rate) {
    const size_t n = a.n_series.size();
    if (n < 2) return 0.0;
    const double peak = *std::max_element(a.n_series.begin(), a.n_series.end());
    if (peak <= 0.0) return 0.0;
    const double cap = 0.6 * peak;
    std::vector<double> t, l;
    for (size_t i = 0; i < n; ++i) {
        const double N = a.n_series[i];
        if (N > 0.0 && N <= cap) {
            t.push_back(a.t_series[i]);
            l.push_back(std::log(N));
        }
    }
    if (t.size() < 2) return 0.0;
    double mt = 0.0, ml = 0.0;
    for (size_t i = 0; i < t.size(); ++i) {
        mt += t[i];
        ml += l[i];
    }
    mt /= static_cast<double>(t.size());
    ml /= static_cast<double>(t.size());
    double stl = 0.0, stt = 0.0;
    for (size_t i = 0; i < t.size(); ++i) {
        const double d = t[i] - mt;
        stl += d * (l[i] - ml);
        stt += d * d;
    }
    if (stt <= 0.0) return 0.0;
    const double slope = stl / stt;
    if (slope <= 0.0) return 0.0;
    return (std::log(2.0) / slope) * biology_rate;
}

// metric_measure (src/sim/accept.cpp:185-206).
int32_t measure(int32_t metric, const astro_stats& fin, const Aggregates& a, astro_world* w,
                const astro_scenario& s, double& out) {
    out = 0.0;
    switch (metric) {
        case ASTRO_METRIC_AWAKE_FRACTION:
            out = fin.n_live > 0 ? static_cast<double>(fin.n_awake) / static_cast<double>(fin.n_live) : 0.0;
            return ASTRO_OK;
        case ASTRO_METRIC_MEDIUM_TEMP_MEAN: out = fin.mean_temp_medium_k; return ASTRO_OK;
        case ASTRO_METRIC_MEDIUM_TEMP_MAX: out = a.max_medium_k; return ASTRO_OK;
        case ASTRO_METRIC_BOIL_EVENT_COUNT: out = static_cast<double>(a.boil_events); return ASTRO_OK;
        case ASTRO_METRIC_POPULATION_LIVE: out = static_cast<double>(fin.n_live); return ASTRO_OK;
        case ASTRO_METRIC_POPULATION_DEAD: out = static_cast<double>(fin.n_dead); return ASTRO_OK;
        case ASTRO_METRIC_MEAN_CHARGE: out = fin.mean_charge; return ASTRO_OK;
        case ASTRO_METRIC_TOTAL_ENERGY_J: out = fin.total_energy_j; return ASTRO_OK;
        case ASTRO_METRIC_MEAN_TAU_TOLERANCE: out = fin.mean_tau_tolerance; return ASTRO_OK;
        case ASTRO_METRIC_RISE_VELOCITY_EMPTY: out = a.rise_velocity_empty; return ASTRO_OK;
        case ASTRO_METRIC_FALL_VELOCITY_FULL: out = a.fall_velocity_full; return ASTRO_OK;
        case ASTRO_METRIC_DOUBLING_TIME_S: {
            double bio = 1.0;
            const int32_t st = astro_world_time(w, nullptr, nullptr, nullptr, &bio);
            out = doubling_time(a, bio);
            return st;
        }
        case ASTRO_METRIC_CHARGE_HEIGHT_CORRELATION:
        case ASTRO_METRIC_CHARGE_DEPTH_CORRELATION: {
            // charge against y (P1), or against depth along the light (P5)
            CellSample c;
            const int32_t st = sample_cells(w, c);
            if (st != ASTRO_OK) return st;
            const size_t n = c.x.size();
            if (n < 2) return ASTRO_OK;
            std::vector<double> charge(n), other(n);
            const double lx = s.lights[0].dir_x, ly = s.lights[0].dir_y;
            for (size_t i = 0; i < n; ++i) {
                charge[i] = c.energy[i] / ASTRO_CELL_ENERGY_MAX;
                other[i] = metric == ASTRO_METRIC_CHARGE_HEIGHT_CORRELATION ? c.y[i] : c.x[i] * lx + c.y[i] * ly;
            }
            out = correlation(charge, other);
            return ASTRO_OK;
        }
        case ASTRO_METRIC_IMPULSE_PER_CYCLE: {
            double imp = 0.0, dis = 0.0;
            const int32_t st = astro_world_flash_audit(w, &imp, &dis);
            out = dis > 0.0 ? imp * ASTRO_C_LIGHT / dis : 0.0;
            return st;
        }
        case ASTRO_METRIC_MAX_TAU_TOLERANCE: {
            // over live Taumoeba (src/sim/accept.cpp:63-77)
            int32_t n = 0, cap = 0;
            int32_t st = astro_taumoeba_counters(w, &n, &cap, nullptr);
            if (st != ASTRO_OK || n <= 0) return st;
            std::vector<uint64_t> id(n), rng(n);
            std::vector<uint32_t> flags(n), gen(n);
            std::vector<double> d64[8];
            std::vector<float> f32[7];
            std::vector<int32_t> target(n);
            for (auto& v : d64) v.resize(n);
            for (auto& v : f32) v.resize(n);
            astro_tau_soa t{id.data(),     flags.data(),  d64[0].data(), d64[1].data(), d64[2].data(),
                            d64[3].data(), d64[4].data(), d64[5].data(), f32[0].data(), f32[1].data(),
                            f32[2].data(), d64[6].data(), d64[7].data(), f32[3].data(), f32[4].data(),
                            gen.data(),    rng.data(),    f32[5].data(), f32[6].data(), target.data()};
            int32_t got = 0;
            if ((st = astro_taumoeba_download(w, &t, n, &got)) != ASTRO_OK) return st;
            for (int32_t i = 0; i < n; ++i)
                if ((flags[i] & ASTRO_FLAG_OCCUPIED) && (flags[i] & ASTRO_FLAG_ALIVE))
                    out = std::max(out, static_cast<double>(f32[4][i]));
            return ASTRO_OK;
        }
        default: return ASTRO_OK;
    }
}

// csv_header / csv_row (tools/headless.cpp:157-173).
void csv_header(std::FILE* f, const astro_scenario& s, int non_canon) {
    std::fprintf(f, "# scenario: %s\n# seed: %llu\n# non_canon_run: %d%s\n", s.id,
                 static_cast<unsigned long long>(s.seed), non_canon, non_canon ? " (CANON LOCK BROKEN)" : " (canon)");
    std::fprintf(f, "# git_describe: astro_hot\n");
    std::fprintf(f,
                 "t_s,n_live,n_dead,n_tau,mean_charge,total_energy_J,mean_temp_cell_K,"
                 "mean_temp_medium_K,co2_total_kg,mean_tolerance,boil_events,divisions,deaths\n");
}
void csv_row(std::FILE* f, const astro_stats& s) {
    std::fprintf(f, "%.6f,%d,%d,%d,%.6f,%.6e,%.4f,%.4f,%.6e,%.6f,%d,%d,%d\n", s.sim_time_s, s.n_live, s.n_dead,
                 s.n_taumoeba, s.mean_charge, s.total_energy_j, s.mean_temp_cell_k, s.mean_temp_medium_k,
                 s.co2_total_kg, s.mean_tau_tolerance, s.boil_events, s.divisions_this_window,
                 s.deaths_this_window);
}

}  // namespace

extern "C" {

int32_t astro_scenario_parse(const char* text, size_t len, astro_scenario* out) {
    if (!out || (!text && len)) return ASTRO_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    jsonc::Node root;
    std::string err;
    if (!jsonc::Reader(std::string_view(text ? text : "", len)).read(root, err) || root.kind != jsonc::Node::OBJ)
        return ASTRO_INVALID_ARGUMENT;
    const int32_t st = read_scenario(root, *out);
    if (st != ASTRO_OK) std::memset(out, 0, sizeof(*out));
    return st;
}

int32_t astro_scenario_load(const char* path, astro_scenario* out) {
    if (!path || !out) return ASTRO_INVALID_ARGUMENT;
    std::ifstream f(path, std::ios::binary);
    if (!f) return ASTRO_INVALID_ARGUMENT;   // as the reference: cannot open
    const std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    return astro_scenario_parse(text.data(), text.size(), out);
}

// scenario_capacities (src/sim/scenario.cpp:327-341).
void astro_scenario_capacities(const astro_scenario* s, int32_t* cell_capacity, int32_t* tau_capacity) {
    int64_t cells = 0, tau = 0;
    if (s)
        for (int i = 0; i < s->population_count; ++i)
            (s->populations[i].kind == ASTRO_ORGANISM_TAUMOEBA ? tau : cells) += s->populations[i].count;
    int64_t cap = cells * 2 + 4096;   // headroom for division
    cap = std::clamp<int64_t>(cap, ASTRO_DEFAULT_CELLS, ASTRO_MAX_CELLS);
    int64_t tcap = tau * 4 + ASTRO_DEFAULT_TAUMOEBA;
    tcap = std::min<int64_t>(tcap, ASTRO_MAX_TAUMOEBA);
    if (cell_capacity) *cell_capacity = static_cast<int32_t>(cap);
    if (tau_capacity) *tau_capacity = static_cast<int32_t>(tcap);
}

// scenario_instantiate (src/sim/scenario.cpp:343-397).
int32_t astro_scenario_instantiate(const astro_scenario* s, const astro_world_desc* base, astro_world** out) {
    if (!s || !out) return ASTRO_INVALID_ARGUMENT;
    *out = nullptr;
    astro_world_desc d;
    if (base) d = *base;
    else astro_world_desc_default(&d);
    d.chamber_w = s->chamber_w;
    d.chamber_h = s->chamber_h;
    d.chamber_d = s->chamber_d;
    astro_scenario_capacities(s, &d.capacity, &d.tau_capacity);
    d.seed = s->seed;
    d.co2_init = s->co2_init;
    d.boundary_x = static_cast<uint8_t>(s->boundary_x);
    d.boundary_y = static_cast<uint8_t>(s->boundary_y);
    d.gravity_axis = static_cast<uint8_t>(s->gravity_axis);
    d.ambient_temp = s->ambient_temp;
    d.store_disposition = static_cast<uint8_t>(s->store_disposition);
    d.thermal_enabled = s->thermal_active ? 1 : 0;
    d.compaction_enabled = s->compaction ? 1 : 0;
    d.tau_compaction_enabled = s->tau_compaction ? 1 : 0;
    astro_world* w = nullptr;
    int32_t st = astro_world_create(&d, &w);
    if (st != ASTRO_OK) return st;

    // T starts at ambient and N2 at zero; fill only what differs
    if (st == ASTRO_OK && s->temp_init != s->ambient_temp)
        st = astro_field_fill(w, ASTRO_FIELD_TEMPERATURE, static_cast<float>(s->temp_init));
    if (st == ASTRO_OK && s->n2_init != 0.0) st = astro_field_fill(w, ASTRO_FIELD_N2, static_cast<float>(s->n2_init));

    // one seed per non-empty population, the first being the scenario's
    uint64_t seed = s->seed;
    for (int i = 0; st == ASTRO_OK && i < s->population_count; ++i) {
        const astro_population_spec& p = s->populations[i];
        if (p.count <= 0) continue;
        if (p.kind == ASTRO_ORGANISM_TAUMOEBA) {
            st = astro_taumoeba_spawn(w, p.count, seed);
            ++seed;
            continue;
        }
        astro_spawn_params sp{};
        sp.count = p.count;
        sp.placement = p.placement;
        sp.place_x = p.place_x;
        sp.place_y = p.place_y;
        sp.place_radius = p.place_radius;
        sp.charge_dist = p.charge_dist;
        sp.charge_a = p.charge_a;
        sp.charge_b = p.charge_b;
        sp.awake = p.awake;
        st = astro_cells_spawn(w, &sp, seed);
        ++seed;
    }
    if (st == ASTRO_OK) {
        const astro_light_source& l = s->lights[0];
        st = l.enabled ? astro_world_set_light(w, l.dir_x, l.dir_y, l.irradiance, s->ambient_irradiance, 1)
                       : astro_world_set_light(w, d.light_dir_x, d.light_dir_y, d.light_irradiance,
                                               s->ambient_irradiance, d.light_enabled);
    }
    if (st == ASTRO_OK) st = astro_world_set_clock_preset(w, s->clock, s->physics_rate, s->biology_rate);
    if (st != ASTRO_OK) {
        astro_world_destroy(w);
        return st;
    }
    *out = w;
    return ASTRO_OK;
}

// scenario_apply_drive (src/sim/scenario.cpp:399-468). World calls that fail
// are reported (the reference discards their status).
int32_t astro_scenario_apply_drive(astro_world* w, const astro_scenario* s) {
    if (!w || !s) return ASTRO_INVALID_ARGUMENT;
    uint64_t tick = 0;
    double t = 0.0, physics = 1.0;
    int32_t st = astro_world_time(w, &tick, &t, &physics, nullptr);
    if (st != ASTRO_OK) return st;
    const double dt = ASTRO_DT_PHYSICS * physics;
    astro_world_desc d;
    if ((st = astro_world_desc_get(w, &d)) != ASTRO_OK) return st;
    // the flash is re-armed each tick by an active stimulus
    bool flash = false;
    for (int i = 0; i < s->drive_count && st == ASTRO_OK; ++i) {
        const astro_stimulus& e = s->drive[i];
        if (!stimulus_on(e, t)) continue;
        const bool open_ended = e.t1_s <= e.t0_s;
        const double frac = open_ended ? 0.0 : (t - e.t0_s) / (e.t1_s - e.t0_s);
        const double strength = e.v0 + (e.v1 - e.v0) * frac;
        // a global brush: a radius far beyond the chamber keeps the whole
        // field in the flat top of the falloff
        const double radius = e.radius > 0.0 ? e.radius : 10.0 * (d.chamber_w + d.chamber_h);
        switch (e.kind) {
            case ASTRO_STIM_HEAT: st = astro_world_apply_brush(w, ASTRO_BRUSH_HEAT, e.x, e.y, radius, strength * dt); break;
            case ASTRO_STIM_CHILL: st = astro_world_apply_brush(w, ASTRO_BRUSH_CHILL, e.x, e.y, radius, strength * dt); break;
            case ASTRO_STIM_INJECT_CO2:
                st = astro_world_apply_brush(w, ASTRO_BRUSH_INJECT_CO2, e.x, e.y, radius, strength * dt);
                break;
            case ASTRO_STIM_INJECT_N2:
                st = astro_world_apply_brush(w, ASTRO_BRUSH_INJECT_N2, e.x, e.y, radius, strength * dt);
                break;
            case ASTRO_STIM_SET_LIGHT:   // an absolute level; direction unchanged
                st = astro_world_set_light(w, d.light_dir_x, d.light_dir_y, static_cast<float>(strength),
                                           d.ambient_irradiance, strength > 0.0 ? 1 : 0);
                if (st == ASTRO_OK) st = astro_world_desc_get(w, &d);
                break;
            case ASTRO_STIM_SET_N2: st = astro_field_fill(w, ASTRO_FIELD_N2, static_cast<float>(strength)); break;
            case ASTRO_STIM_SEED_CELLS: {
                // keep the astrophage count at `strength`, spawning the deficit
                // with the first astrophage population's spec
                astro_counters c{};
                if ((st = astro_world_get_counters(w, &c)) != ASTRO_OK) break;
                const int32_t deficit = static_cast<int32_t>(strength) - c.count;
                if (deficit <= 0) break;
                for (int p = 0; p < s->population_count; ++p) {
                    const astro_population_spec& ps = s->populations[p];
                    if (ps.kind != ASTRO_ORGANISM_ASTROPHAGE) continue;
                    astro_spawn_params sp{};
                    sp.count = deficit;
                    sp.placement = ps.placement;
                    sp.place_x = ps.place_x;
                    sp.place_y = ps.place_y;
                    sp.place_radius = ps.place_radius;
                    sp.charge_dist = ps.charge_dist;
                    sp.charge_a = ps.charge_a;
                    sp.charge_b = ps.charge_b;
                    sp.awake = ps.awake;
                    st = astro_cells_spawn(w, &sp, d.seed + tick);
                    break;
                }
                break;
            }
            case ASTRO_STIM_FLASH: flash = true; break;
            default: st = ASTRO_INVALID_ARGUMENT; break;
        }
    }
    if (st == ASTRO_OK) st = astro_world_set_flash(w, flash ? 1 : 0);
    return st;
}

// accept_eval (src/sim/accept.cpp:208-221).
int32_t astro_accept_eval(const astro_accept_check* c, double m) {
    if (!c) return 0;
    switch (c->op) {
        case ASTRO_OP_EQ: return m == c->value;
        case ASTRO_OP_NE: return m != c->value;
        case ASTRO_OP_LT: return m < c->value;
        case ASTRO_OP_LE: return m <= c->value;
        case ASTRO_OP_GT: return m > c->value;
        case ASTRO_OP_GE: return m >= c->value;
        case ASTRO_OP_APPROX: {
            const double tol = c->tol_absolute ? c->tol : c->tol * std::fabs(c->value);
            return std::fabs(m - c->value) <= tol;
        }
        default: return 0;
    }
}

const char* astro_metric_name(int32_t metric) {
    for (const Named& e : METRICS)
        if (e.value == metric) return e.name;
    return "?";
}

// run_scenario_assert (tools/headless.cpp:179-250). Ticks in which no
// stimulus can be active are stepped in batches up to the next stats sample,
// so they replay the world's captured graph.
int32_t astro_scenario_run_assert(const astro_scenario* s, astro_world* w, int64_t ticks_override,
                                  const char* csv_path, astro_accept_result* results, int32_t cap,
                                  astro_assert_summary* summary) {
    if (!s || !w || (cap > 0 && !results)) return ASTRO_INVALID_ARGUMENT;
    astro_assert_summary sum{};
    if (s->accept_count == 0) {   // no objective: accepted without running
        sum.all_pass = 1;
        if (summary) *summary = sum;
        return ASTRO_OK;
    }
    double horizon = s->run_duration_s;
    if (horizon <= 0.0) {
        for (int i = 0; i < s->accept_count; ++i) horizon = std::max(horizon, s->accept[i].after_s);
        for (int i = 0; i < s->drive_count; ++i)
            horizon = std::max(horizon, s->drive[i].t1_s > s->drive[i].t0_s ? s->drive[i].t1_s : s->drive[i].t0_s);
    }
    if (horizon <= 0.0) horizon = 60.0;
    double physics = 1.0;
    int32_t st = astro_world_time(w, nullptr, nullptr, &physics, nullptr);
    if (st != ASTRO_OK) return st;
    const double dt = ASTRO_DT_PHYSICS * physics;
    int64_t ticks = ticks_override > 0 ? ticks_override : static_cast<int64_t>(std::ceil(horizon / dt));
    if (ticks < 1) ticks = 1;
    const int64_t stride = ticks > 2000 ? ticks / 2000 : 1;
    const double settle_s = 2.0, interval_s = 4.0;   // velocity capture windows

    const Needs needs = needs_of(*s);
    Aggregates agg;
    astro_stats stats{};
    std::FILE* csv = csv_path ? std::fopen(csv_path, "w") : nullptr;
    if (csv_path && !csv) return ASTRO_FILE_IO_ERROR;
    bool header = false;
    auto take_sample = [&]() -> int32_t {
        int32_t rc = astro_world_stats(w, &stats);
        if (rc != ASTRO_OK) return rc;
        if (csv && !header) {
            csv_header(csv, *s, stats.non_canon_run);
            header = true;
        }
        rc = sample(agg, stats, w, needs, settle_s, interval_s, s->chamber_h);
        if (rc == ASTRO_OK && csv) csv_row(csv, stats);
        ++sum.samples;
        return rc;
    };

    int64_t t = 0;   // ticks done; iteration t applies the drive, steps, and samples when t % stride == 0
    while (st == ASTRO_OK && t < ticks) {
        double now = 0.0;
        if ((st = astro_world_time(w, nullptr, &now, nullptr, nullptr)) != ASTRO_OK) break;
        // ticks that certainly start outside every stimulus window
        int64_t quiet = ticks - t;
        for (int i = 0; i < s->drive_count; ++i) {
            const astro_stimulus& e = s->drive[i];
            const bool open_ended = e.t1_s <= e.t0_s;
            if (stimulus_on(e, now)) {
                quiet = 0;
                break;
            }
            if (!open_ended && now >= e.t1_s) continue;   // window over
            const double ahead = std::floor((e.t0_s - now) / dt) - 1.0;
            quiet = std::min<int64_t>(quiet, ahead > 0.0 ? static_cast<int64_t>(ahead) : 0);
        }
        if (quiet <= 0) {
            st = astro_scenario_apply_drive(w, s);
            if (st == ASTRO_OK) st = astro_world_step(w, 1);
            if (st == ASTRO_OK && t % stride == 0) st = take_sample();
            ++t;
            continue;
        }
        // no stimulus acts in these ticks: the drive only disarms the flash
        if ((st = astro_scenario_apply_drive(w, s)) != ASTRO_OK) break;
        const int64_t next_sample = t % stride == 0 ? t : (t / stride + 1) * stride;
        const int64_t k = std::min<int64_t>({next_sample - t + 1, ticks - t, quiet});
        st = astro_world_step(w, static_cast<int32_t>(k));
        t += k;
        if (st == ASTRO_OK && (t - 1) % stride == 0) st = take_sample();
    }
    if (st == ASTRO_OK) st = astro_world_sync(w);
    if (st == ASTRO_OK) st = take_sample();   // the final state
    if (csv) std::fclose(csv);
    if (st != ASTRO_OK) return st;

    sum.ticks = t;
    sum.sim_time_s = stats.sim_time_s;
    sum.checks = s->accept_count;
    sum.all_pass = 1;
    for (int i = 0; i < s->accept_count; ++i) {
        const astro_accept_check& c = s->accept[i];
        double m = 0.0;
        if ((st = measure(c.metric, stats, agg, w, *s, m)) != ASTRO_OK) return st;
        const int32_t pass = astro_accept_eval(&c, m);
        sum.passed += pass;
        sum.all_pass &= pass;
        if (i < cap) {
            astro_accept_result& r = results[i];
            r.metric = c.metric;
            r.op = c.op;
            r.measured = m;
            r.value = c.value;
            r.tol = c.tol;
            r.after_s = c.after_s;
            r.tol_absolute = c.tol_absolute;
            r.pass = pass;
        }
    }
    if (summary) *summary = sum;
    return ASTRO_OK;
}

}  // extern "C"
