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
#include <cstddef>
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

// ---- the scenario schema -----------------------------------------------------
// The JSON vocabulary, defaults and caps are the reference's
// (src/sim/scenario.cpp:139-318, contracts/scenario_v3.h); the reader here is
// declarative: each object of the document is a table of Fields, read the same
// way whether the object is present or not (absent members, and an absent
// object, take the defaults).
enum class Ty { F64, F32, I32, U64, FLAG, TEXT, NAME };

struct Field {
    const char* keys;     // "key" or "key|alternative|...": the first present wins
    Ty ty;
    size_t off;           // byte offset in the target struct
    double def = 0.0;     // F64 / F32 / I32 / U64 / FLAG default
    const char* def_name = "";             // TEXT / NAME default
    const Named* names = nullptr;          // NAME vocabulary
    size_t n_names = 0;
    size_t cap = 0;                        // TEXT capacity
};

template <size_t N>
Field name_field(const char* key, size_t off, const Named (&v)[N], const char* def) {
    return Field{key, Ty::NAME, off, 0.0, def, v, N, 0};
}
Field text_field(const char* key, size_t off, size_t cap) { return Field{key, Ty::TEXT, off, 0.0, "", nullptr, 0, cap}; }

// The member an alternative-key list names, or null.
const jsonc::Node* member(const jsonc::Node* obj, const char* keys) {
    if (!obj) return nullptr;
    const char* k = keys;
    while (*k) {
        const char* bar = std::strchr(k, '|');
        const size_t len = bar ? static_cast<size_t>(bar - k) : std::strlen(k);
        if (const jsonc::Node* m = obj->at(std::string_view(k, len))) return m;
        k += len + (bar ? 1 : 0);
    }
    return nullptr;
}

// Reads every field of `table` from `obj` (null: all defaults) into `base`.
// A NAME outside its vocabulary takes the default, unless `strict`, which
// makes it an error (a dropped stimulus or check would change the run).
template <size_t N>
bool read_fields(const jsonc::Node* obj, void* base, const Field (&table)[N], bool strict = false) {
    char* b = static_cast<char*>(base);
    for (const Field& f : table) {
        const jsonc::Node* m = member(obj, f.keys);
        const double num = m ? m->as_num(f.def) : f.def;
        switch (f.ty) {
            case Ty::F64: *reinterpret_cast<double*>(b + f.off) = num; break;
            case Ty::F32: *reinterpret_cast<float*>(b + f.off) = static_cast<float>(num); break;
            case Ty::I32: *reinterpret_cast<int32_t*>(b + f.off) = static_cast<int32_t>(num); break;
            case Ty::U64: *reinterpret_cast<uint64_t*>(b + f.off) = static_cast<uint64_t>(num); break;
            case Ty::FLAG:
                *reinterpret_cast<int32_t*>(b + f.off) = (m ? m->as_bool(f.def != 0.0) : f.def != 0.0) ? 1 : 0;
                break;
            case Ty::TEXT: copy_text(b + f.off, f.cap, m ? m->as_str("") : std::string()); break;
            case Ty::NAME: {
                const std::string text = m ? m->as_str(f.def_name) : std::string(f.def_name);
                int32_t v = -1;
                for (size_t i = 0; i < f.n_names; ++i)
                    if (text == f.names[i].name) v = f.names[i].value;
                if (v < 0) {
                    if (strict) return false;
                    for (size_t i = 0; i < f.n_names; ++i)
                        if (std::string_view(f.def_name) == f.names[i].name) v = f.names[i].value;
                }
                *reinterpret_cast<int32_t*>(b + f.off) = v;
                break;
            }
        }
    }
    return true;
}

// Reads up to `cap` elements of the list `key` of `obj`, each with `table`,
// into consecutive structs of type T; returns the count, or -1 on a strict
// name outside its vocabulary.
template <class T, size_t N>
int read_list(const jsonc::Node& obj, const char* key, T* out, int cap, const Field (&table)[N], bool strict = false) {
    int n = 0;
    for (const jsonc::Node& e : obj.list(key)) {
        if (n >= cap) break;
        if (!read_fields(&e, &out[n], table, strict)) return -1;
        ++n;
    }
    return n;
}

const Named GRAVITY_AXES[] = {{"y", ASTRO_GRAVITY_Y}, {"z", ASTRO_GRAVITY_Z}};

#define AT(field) offsetof(astro_scenario, field)
const Field TOP[] = {
    {"schema", Ty::I32, AT(schema), 1.0},
    text_field("id", AT(id), sizeof(astro_scenario::id)),
    text_field("title", AT(title), sizeof(astro_scenario::title)),
    text_field("blurb", AT(blurb), sizeof(astro_scenario::blurb)),
    {"seed", Ty::U64, AT(seed), 20260802.0},
    {"compaction", Ty::FLAG, AT(compaction), 0.0},
    {"tau_compaction", Ty::FLAG, AT(tau_compaction), 0.0},
};
const Field CHAMBER[] = {
    {"w", Ty::F64, AT(chamber_w), ASTRO_CHAMBER_W},
    {"h", Ty::F64, AT(chamber_h), ASTRO_CHAMBER_H},
    {"d", Ty::F64, AT(chamber_d), ASTRO_CHAMBER_D},
    name_field("boundary_x", AT(boundary_x), BOUNDARIES, "reflecting"),
    name_field("boundary_y", AT(boundary_y), BOUNDARIES, "reflecting"),
};
const Field MEDIUM[] = {
    {"temp_init", Ty::F64, AT(temp_init), ASTRO_AMBIENT_TEMP_DEFAULT},
    {"co2_init", Ty::F64, AT(co2_init), 0.0},
    {"n2_init", Ty::F64, AT(n2_init), 0.0},
    {"ambient_temp", Ty::F64, AT(ambient_temp), ASTRO_AMBIENT_TEMP_DEFAULT},
    name_field("thermal_bc", AT(thermal_bc), THERMAL_BCS, "robin"),
    name_field("density_model", AT(density_model), DENSITIES, "canon-mass"),
    name_field("store_disposition", AT(store_disposition), DISPOSITIONS, "void"),
    name_field("gravity_axis", AT(gravity_axis), GRAVITY_AXES, "y"),
    {"thermal_active", Ty::FLAG, AT(thermal_active), 1.0},
};
const Field LIGHT[] = {{"ambient", Ty::F32, AT(ambient_irradiance), 0.0}};
const Field CLOCK[] = {
    name_field("preset", AT(clock), CLOCKS, "realtime"),
    {"physics_rate", Ty::F64, AT(physics_rate), 1.0},
    {"biology_rate", Ty::F64, AT(biology_rate), 1.0},
};
const Field OBJECTIVE[] = {
    text_field("text", AT(objective_text), sizeof(astro_scenario::objective_text)),
    {"duration_s", Ty::F64, AT(run_duration_s), 0.0},
};
#undef AT

#define AT(field) offsetof(astro_light_source, field)
const Field SOURCE[] = {
    {"dir_x", Ty::F32, AT(dir_x), 1.0}, {"dir_y", Ty::F32, AT(dir_y), 0.0},
    {"irradiance", Ty::F32, AT(irradiance), 0.0}, {"wavelength", Ty::F32, AT(wavelength), 0.0},
};
#undef AT
#define AT(field) offsetof(astro_population_spec, field)
const Field POPULATION[] = {
    name_field("kind", AT(kind), ORGANISMS, "astrophage"),
    {"count", Ty::I32, AT(count), 0.0},
    name_field("placement", AT(placement), PLACEMENTS, "uniform"),
    {"place_x", Ty::F64, AT(place_x), 0.0},
    {"place_y", Ty::F64, AT(place_y), 0.0},
    {"place_radius", Ty::F64, AT(place_radius), 0.0},
    {"awake", Ty::FLAG, AT(awake), 0.0},
};
const Field CHARGE[] = {
    name_field("dist", AT(charge_dist), DISTRIBUTIONS, "constant"),
    {"value|a", Ty::F64, AT(charge_a), 0.0},
    {"b", Ty::F64, AT(charge_b), 0.0},
};
#undef AT
#define AT(field) offsetof(astro_stimulus, field)
const Field STIMULUS[] = {
    {"kind", Ty::NAME, AT(kind), 0.0, "", STIMULI, std::size(STIMULI), 0},
    {"t0_s", Ty::F64, AT(t0_s), 0.0}, {"t1_s", Ty::F64, AT(t1_s), 0.0},
    {"x", Ty::F64, AT(x), 0.0}, {"y", Ty::F64, AT(y), 0.0}, {"radius", Ty::F64, AT(radius), 0.0},
    {"v0|value", Ty::F64, AT(v0), 0.0},
};
#undef AT
#define AT(field) offsetof(astro_accept_check, field)
const Field CHECK[] = {
    {"metric", Ty::NAME, AT(metric), 0.0, "", METRICS, std::size(METRICS), 0},
    name_field("op", AT(op), OPS, "=="),
    {"value", Ty::F64, AT(value), 0.0},
    {"tol", Ty::F64, AT(tol), 0.0},
    {"after_s", Ty::F64, AT(after_s), 0.0},
    {"tol_absolute", Ty::FLAG, AT(tol_absolute), 0.0},
};
#undef AT

// scenario_parse (src/sim/scenario.cpp:139-318) onto a zeroed scenario.
int32_t read_scenario(const jsonc::Node& root, astro_scenario& o) {
    read_fields(&root, &o, TOP);
    read_fields(root.at("chamber"), &o, CHAMBER);
    read_fields(root.at("medium"), &o, MEDIUM);
    read_fields(root.at("clock"), &o, CLOCK);
    const jsonc::Node* light = root.at("light");
    read_fields(light, &o, LIGHT);
    if (light) {
        const int n = read_list(*light, "sources", o.lights, ASTRO_SCENARIO_MAX_LIGHTS, SOURCE);
        for (int i = 0; i < n; ++i) o.lights[i].enabled = 1;
    }
    o.population_count = read_list(root, "populations", o.populations, ASTRO_SCENARIO_MAX_POPULATIONS, POPULATION);
    for (int i = 0; i < o.population_count; ++i) read_fields(root.list("populations")[i].at("charge"), &o.populations[i], CHARGE);
    for (const jsonc::Node& t : root.list("tools"))
        if (t.kind == jsonc::Node::STR) o.tools |= static_cast<uint32_t>(pick(TOOLS, t.s, 0));
    o.drive_count = read_list(root, "drive", o.drive, ASTRO_SCENARIO_MAX_STIMULI, STIMULUS, true);
    if (o.drive_count < 0) return ASTRO_INVALID_ARGUMENT;
    // a window's end value defaults to its start value
    for (int i = 0; i < o.drive_count; ++i) {
        const jsonc::Node* v1 = member(&root.list("drive")[i], "v1|value");
        o.drive[i].v1 = v1 ? v1->as_num(o.drive[i].v0) : o.drive[i].v0;
    }
    // Overrides are carried, not applied, as in the reference.
    if (const jsonc::Node* po = root.at("param_overrides"); po && po->kind == jsonc::Node::OBJ)
        for (const auto& [key, value] : po->members) {
            if (o.override_count >= ASTRO_SCENARIO_MAX_OVERRIDES) break;
            astro_param_override& d = o.param_overrides[o.override_count++];
            copy_text(d.key, sizeof(d.key), key);
            d.value = value.as_num(0.0);
        }
    const jsonc::Node* objective = root.at("objective");
    read_fields(objective, &o, OBJECTIVE);
    if (objective) {
        o.accept_count = read_list(*objective, "accept", o.accept, ASTRO_SCENARIO_MAX_ACCEPT, CHECK, true);
        if (o.accept_count < 0) return ASTRO_INVALID_ARGUMENT;
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

// Least squares of y on x over pairs added one at a time: running means and
// co-moments updated as in Welford's method, so no second pass and no
// cancellation between large sums. Gives the three fits the acceptance
// metrics use (src/sim/accept.cpp): Pearson's r (0 when either variance is
// 0 or fewer than two pairs), and a line y = a + b x (b = 0 when x does not
// vary).
class LineFit {
public:
    void add(double x, double y) {
        n_ += 1.0;
        const double dx = x - mx_;
        mx_ += dx / n_;
        const double dy = y - my_;
        my_ += dy / n_;
        sxx_ += dx * (x - mx_);
        syy_ += dy * (y - my_);
        sxy_ += dx * (y - my_);
    }
    size_t count() const { return static_cast<size_t>(n_); }
    double r() const { return (n_ >= 2.0 && sxx_ > 0.0 && syy_ > 0.0) ? sxy_ / std::sqrt(sxx_ * syy_) : 0.0; }
    double slope() const { return sxx_ > 0.0 ? sxy_ / sxx_ : 0.0; }
    double intercept() const { return my_ - slope() * mx_; }

private:
    double n_ = 0.0, mx_ = 0.0, my_ = 0.0, sxx_ = 0.0, syy_ = 0.0, sxy_ = 0.0;
};

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

// What a run accumulates for its checks between stats samples (the
// reference's RunAggregates and aggregates_sample, src/sim/accept.h:31-53,
// src/sim/accept.cpp:113-183): the worst medium temperature and the boil
// events; the live population over time, when a doubling time is checked;
// and, when a settling speed is checked, two position snapshots `interval_s`
// apart taken after `settle_s`, fitted to speed = a + b * charge.
class RunRecord {
public:
    RunRecord(const astro_scenario& s, double settle_s, double interval_s)
        : settle_s_(settle_s), interval_s_(interval_s), half_h_(0.5 * s.chamber_h) {
        for (int i = 0; i < s.accept_count; ++i) {
            const int32_t m = s.accept[i].metric;
            track_speed_ |= m == ASTRO_METRIC_RISE_VELOCITY_EMPTY || m == ASTRO_METRIC_FALL_VELOCITY_FULL;
            track_growth_ |= m == ASTRO_METRIC_DOUBLING_TIME_S;
        }
    }

    int32_t observe(const astro_stats& st, astro_world* w) {
        boil_events_ += st.boil_events;
        hottest_k_ = std::max(hottest_k_, st.max_temp_medium_k);
        if (track_growth_) growth_.emplace_back(st.sim_time_s, static_cast<double>(st.n_live));
        if (!track_speed_ || speed_done_) return ASTRO_OK;
        const bool first = !have_first_ && st.sim_time_s >= settle_s_;
        const bool second = have_first_ && st.sim_time_s >= t_first_ + interval_s_;
        if (!first && !second) return ASTRO_OK;
        CellSample c;
        const int32_t rc = sample_cells(w, c);
        if (rc != ASTRO_OK || c.x.empty()) return rc;
        if (first) {
            y_first_ = c.y;
            t_first_ = st.sim_time_s;
            charge_first_.clear();
            for (double e : c.energy) charge_first_.push_back(e / ASTRO_CELL_ENERGY_MAX);
            have_first_ = true;
            return ASTRO_OK;
        }
        fit_speed(c, st.sim_time_s - t_first_);
        speed_done_ = true;
        return ASTRO_OK;
    }

    int64_t boil_events() const { return boil_events_; }
    double hottest_k() const { return hottest_k_; }
    double empty_speed() const { return empty_speed_; }
    double full_speed() const { return full_speed_; }

    // ln N = c + t ln2 / T_d over the samples below 60 % of the peak (past
    // the peak the medium is exhausted), T_d in culture time: the simulated
    // time scaled by the biology rate (src/sim/accept.cpp:85-108).
    double doubling_time(double biology_rate) const {
        double peak = 0.0;
        for (const auto& [t, n] : growth_) peak = std::max(peak, n);
        LineFit fit;
        for (const auto& [t, n] : growth_)
            if (n > 0.0 && n <= 0.6 * peak) fit.add(t, std::log(n));
        const double rate = fit.count() >= 2 ? fit.slope() : 0.0;
        return rate > 0.0 ? biology_rate * std::log(2.0) / rate : 0.0;
    }

private:
    // Downward speed of each cell away from the y walls (3 diameters) against
    // its charge at the first snapshot: a = the empty cell's rise, a + b the
    // full cell's fall.
    void fit_speed(const CellSample& c, double dt) {
        if (!(dt > 0.0)) return;
        LineFit fit;
        const size_t m = std::min(c.y.size(), y_first_.size());
        for (size_t i = 0; i < m; ++i)
            if (std::fabs(c.y[i]) <= half_h_ - 3.0 * ASTRO_CELL_DIAMETER)
                fit.add(charge_first_[i], (y_first_[i] - c.y[i]) / dt);
        if (fit.count() < 2) return;
        empty_speed_ = fit.intercept();
        full_speed_ = fit.intercept() + fit.slope();
    }

    double settle_s_, interval_s_, half_h_;
    bool track_speed_ = false, track_growth_ = false;
    int64_t boil_events_ = 0;
    double hottest_k_ = 0.0;
    std::vector<std::pair<double, double>> growth_;
    bool have_first_ = false, speed_done_ = false;
    double t_first_ = 0.0;
    std::vector<double> y_first_, charge_first_;
    double empty_speed_ = 0.0, full_speed_ = 0.0;
};

// Pearson r of charge against height (y), or against depth along the first
// light source (P1 and P5 of the reference's acceptance metrics).
int32_t charge_correlation(astro_world* w, bool along_light, const astro_scenario& s, double& out) {
    CellSample c;
    const int32_t st = sample_cells(w, c);
    if (st != ASTRO_OK) return st;
    LineFit fit;
    for (size_t i = 0; i < c.x.size(); ++i) {
        const double pos = along_light ? c.x[i] * s.lights[0].dir_x + c.y[i] * s.lights[0].dir_y : c.y[i];
        fit.add(c.energy[i] / ASTRO_CELL_ENERGY_MAX, pos);
    }
    out = fit.r();
    return ASTRO_OK;
}

// The largest N2 tolerance among live Taumoeba.
int32_t max_tolerance(astro_world* w, double& out) {
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
    const uint32_t live = ASTRO_FLAG_OCCUPIED | ASTRO_FLAG_ALIVE;
    for (int32_t i = 0; i < n; ++i)
        if ((flags[i] & live) == live) out = std::max(out, static_cast<double>(f32[4][i]));
    return ASTRO_OK;
}

// A check's measured value: from the final stats, from what the run
// recorded, or from the world's final state.
int32_t measure(int32_t metric, const astro_stats& fin, const RunRecord& rec, astro_world* w,
                const astro_scenario& s, double& out) {
    out = 0.0;
    const double live = static_cast<double>(fin.n_live);
    switch (metric) {
        case ASTRO_METRIC_AWAKE_FRACTION: out = live > 0.0 ? fin.n_awake / live : 0.0; break;
        case ASTRO_METRIC_MEDIUM_TEMP_MEAN: out = fin.mean_temp_medium_k; break;
        case ASTRO_METRIC_MEDIUM_TEMP_MAX: out = rec.hottest_k(); break;
        case ASTRO_METRIC_BOIL_EVENT_COUNT: out = static_cast<double>(rec.boil_events()); break;
        case ASTRO_METRIC_POPULATION_LIVE: out = live; break;
        case ASTRO_METRIC_POPULATION_DEAD: out = static_cast<double>(fin.n_dead); break;
        case ASTRO_METRIC_MEAN_CHARGE: out = fin.mean_charge; break;
        case ASTRO_METRIC_TOTAL_ENERGY_J: out = fin.total_energy_j; break;
        case ASTRO_METRIC_MEAN_TAU_TOLERANCE: out = fin.mean_tau_tolerance; break;
        case ASTRO_METRIC_RISE_VELOCITY_EMPTY: out = rec.empty_speed(); break;
        case ASTRO_METRIC_FALL_VELOCITY_FULL: out = rec.full_speed(); break;
        case ASTRO_METRIC_MAX_TAU_TOLERANCE: return max_tolerance(w, out);
        case ASTRO_METRIC_CHARGE_HEIGHT_CORRELATION: return charge_correlation(w, false, s, out);
        case ASTRO_METRIC_CHARGE_DEPTH_CORRELATION: return charge_correlation(w, true, s, out);
        case ASTRO_METRIC_DOUBLING_TIME_S: {
            double bio = 1.0;
            const int32_t st = astro_world_time(w, nullptr, nullptr, nullptr, &bio);
            out = rec.doubling_time(bio);
            return st;
        }
        case ASTRO_METRIC_IMPULSE_PER_CYCLE: {
            // |impulse| c / discharged energy: 1 for a perfect photon rocket
            double impulse = 0.0, energy = 0.0;
            const int32_t st = astro_world_flash_audit(w, &impulse, &energy);
            out = energy > 0.0 ? impulse * ASTRO_C_LIGHT / energy : 0.0;
            return st;
        }
        default: break;
    }
    return ASTRO_OK;
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

    RunRecord rec(*s, settle_s, interval_s);
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
        rc = rec.observe(stats, w);
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
        if ((st = measure(c.metric, stats, rec, w, *s, m)) != ASTRO_OK) return st;
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
