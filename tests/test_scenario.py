"""The scenario layer on the CPU: the library's C++ reader against the Python
reading of the reference's scenario files (tests/golden/scenarios/), its
errors, capacities and accept evaluation, and the oracle's brush and clock
presets against what the mathematics and the canon table fix. Parsing is
host code: no GPU call is made."""
import ctypes as C
import math

import numpy as np
import pytest

import scenario_py as sp
from conftest import ROOT, ah

SCENARIOS = sorted((ROOT / "tests" / "golden" / "scenarios").glob("*.json"))
NAMES = [p.stem for p in SCENARIOS]


@pytest.fixture(scope="module")
def lib():
    if not ah.LIB_PATH.exists():
        pytest.skip("libastro_hot.so not built")
    return ah.Binding(ah.LIB_PATH, "astro_", ah.PRODUCT_ONLY)


def test_all_reference_scenarios_present():
    assert {"bloom", "demo-herd", "first-light", "komorov", "sandbox", "shadow-garden", "spin-drive-face",
            "taumoeba", "three-percent-line"} <= set(NAMES)


@pytest.mark.parametrize("name", NAMES)
def test_parse_matches_python_reading(lib, name):
    text = (ROOT / "tests" / "golden" / "scenarios" / f"{name}.json").read_text()
    got = ah.scenario_parse(lib, text)
    want = sp.read(text)
    assert ah.scenario_load(lib, ROOT / "tests" / "golden" / "scenarios" / f"{name}.json").id == got.id
    for k in ("schema", "seed", "chamber_w", "chamber_h", "chamber_d", "boundary_x", "boundary_y", "temp_init",
              "co2_init", "n2_init", "ambient_temp", "thermal_bc", "density_model", "store_disposition",
              "gravity_axis", "clock", "physics_rate", "biology_rate", "tools", "run_duration_s"):
        assert getattr(got, k) == want[k], k
    assert got.id.decode() == want["id"]
    assert (bool(got.thermal_active), bool(got.compaction), bool(got.tau_compaction)) == \
           (want["thermal_active"], want["compaction"], want["tau_compaction"])
    assert got.ambient_irradiance == np.float32(want["ambient_irradiance"])
    assert got.population_count == len(want["populations"])
    for i, p in enumerate(want["populations"]):
        g = got.populations[i]
        for k in ("kind", "count", "placement", "place_x", "place_y", "place_radius", "charge_dist", "charge_a",
                  "charge_b"):
            assert getattr(g, k) == p[k], (i, k)
        assert bool(g.awake) == p["awake"]
    for i, l in enumerate(want["lights"]):
        g = got.lights[i]
        assert g.enabled == 1
        for k in ("dir_x", "dir_y", "irradiance", "wavelength"):
            assert getattr(g, k) == np.float32(l[k]), (i, k)
    assert not any(got.lights[i].enabled for i in range(len(want["lights"]), 8))
    assert got.drive_count == len(want["drive"])
    for i, e in enumerate(want["drive"]):
        for k in ("kind", "t0_s", "t1_s", "x", "y", "radius", "v0", "v1"):
            assert getattr(got.drive[i], k) == e[k], (i, k)
    assert got.accept_count == len(want["accept"])
    for i, c in enumerate(want["accept"]):
        g = got.accept[i]
        for k in ("metric", "op", "value", "tol", "after_s"):
            assert getattr(g, k) == c[k], (i, k)
        assert bool(g.tol_absolute) == c["tol_absolute"]
    assert got.override_count == len(want["overrides"])


def test_parse_errors(lib, tmp_path):
    base = (ROOT / "tests" / "golden" / "scenarios" / "first-light.json").read_text()
    bad_stim = base.replace('"kind": "heat"', '"kind": "teleport"')
    bad_metric = base.replace('"awake_fraction"', '"happiness"')
    for text in (bad_stim, bad_metric, "[1, 2]", "{ \"id\": ", "", "not json"):
        with pytest.raises(ah.AstroError) as e:
            ah.scenario_parse(lib, text)
        assert e.value.status == ah.INVALID_ARGUMENT
    with pytest.raises(ah.AstroError) as e:
        ah.scenario_load(lib, tmp_path / "missing.json")
    assert e.value.status == ah.INVALID_ARGUMENT
    # unknown enum strings fall back to the documented defaults
    sc = ah.scenario_parse(lib, '{"chamber": {"boundary_x": "wobbly"}, "clock": {"preset": "warp"},'
                                ' /* remark */ "medium": {"thermal_bc": "?"}} // trailing')
    assert (sc.boundary_x, sc.clock, sc.thermal_bc, sc.thermal_active) == (0, ah.CLOCK_REALTIME, 2, 1)
    assert (sc.chamber_w, sc.seed, sc.temp_init) == (0.004, 20260802, 293.14999999999998)


def test_strings_and_numbers(lib):
    sc = ah.scenario_parse(lib, r'{"id": "a\"b\\cA", "title": "x//y /* z */", "seed": 1e3,'
                                r' "chamber": {"w": -2.5E-3}}')
    assert sc.id.decode() == 'a"b\\cA' and sc.title.decode() == "x//y /* z */"
    assert sc.seed == 1000 and sc.chamber_w == -2.5e-3


def test_capacities(lib):
    def caps(pops):
        sc = ah.Scenario()
        for i, (kind, n) in enumerate(pops):
            sc.populations[i].kind, sc.populations[i].count = kind, n
        sc.population_count = len(pops)
        c, t = C.c_int32(), C.c_int32()
        lib.scenario_capacities(C.byref(sc), C.byref(c), C.byref(t))
        return c.value, t.value
    # headroom 2n + 4096, clamped to [DEFAULT_CELLS, MAX_CELLS]; Taumoeba 4n + 256 up to 65536
    assert caps([]) == (25000, 256)
    assert caps([(0, 12000)]) == (28096, 256)
    assert caps([(0, 24000), (1, 1024)]) == (52096, 4352)
    assert caps([(0, 1_500_000), (1, 50_000)]) == (2_000_000, 65536)


def test_accept_eval_and_names(lib):
    c = ah.AcceptCheck(metric=0, op=ah.OPS.index("~="), value=100.0, tol=0.05)
    assert lib.accept_eval(C.byref(c), 104.9) and not lib.accept_eval(C.byref(c), 105.1)
    c.tol_absolute, c.tol = 1, 0.5
    assert lib.accept_eval(C.byref(c), 100.5) and not lib.accept_eval(C.byref(c), 100.6)
    for op, yes, no in (("==", 1.0, 1.5), ("!=", 2.0, 1.0), ("<", 0.5, 1.0), ("<=", 1.0, 1.1), (">", 1.5, 1.0),
                        (">=", 1.0, 0.9)):
        k = ah.AcceptCheck(metric=0, op=ah.OPS.index(op), value=1.0)
        assert lib.accept_eval(C.byref(k), yes) == 1 and lib.accept_eval(C.byref(k), no) == 0, op
    assert [lib.metric_name(i).decode() for i in range(16)] == ah.METRICS
    assert lib.metric_name(99) == b"?"


# ---- the oracle's stimuli -------------------------------------------------------
def test_oracle_brush_closed_form(oracle):
    """The smooth disc: the centre voxel gains amount * (1 - (r/R)^2)^2, the
    total over the disc approaches amount * pi R^2 / 3 / (dx dy), nothing
    outside R changes, and chill is heat's negative."""
    grid = {0: (200, 160, 1), 1: (64, 64, 1), 2: (32, 32, 1), 3: (64, 64, 1)}
    with ah.World(oracle, capacity=10, grid=grid) as w:
        T0 = w.field(ah.TEMPERATURE).astype(np.float64)
        (nx, ny, _), (dx, dy, _) = w.field_dims(ah.TEMPERATURE)
        cx, cy, R, amount = 3.1e-4, -2.2e-4, 9.0e-4, 2.5
        w.brush(ah.BRUSH_HEAT, cx, cy, R, amount)
        inc = w.field(ah.TEMPERATURE).astype(np.float64) - T0
        want = sp.brush_increment(inc.shape, dx, dy, cx, cy, R, amount)
        assert np.all(inc[want == 0.0] == 0.0)
        assert np.abs(inc - want).max() < 5e-5          # float rounding of T ~ 293 K plus the weight
        total = inc.sum() * dx * dy
        assert abs(total - amount * math.pi * R * R / 3.0) / (amount * math.pi * R * R / 3.0) < 0.01
        w.brush(ah.BRUSH_CHILL, cx, cy, R, amount)
        back = w.field(ah.TEMPERATURE).astype(np.float64)
        assert np.abs(back - T0).max() < 1e-4
        w.brush(ah.BRUSH_HEAT, 0.0, 0.0, 0.0, 5.0)   # radius 0: nothing
        assert np.array_equal(w.field(ah.TEMPERATURE), back.astype(np.float32))


def test_oracle_brush_3d_columns_and_fill(oracle):
    grid = {0: (24, 20, 6), 1: (12, 10, 3), 2: (-1, 0, 0), 3: (24, 20, 6)}
    with ah.World(oracle, capacity=10, chamber_w=5e-4, chamber_h=4e-4, chamber_d=1e-4, grid=grid) as w:
        w.fill(ah.CO2, 0.25)
        assert np.all(w.field(ah.CO2) == np.float32(0.25))
        w.brush(ah.BRUSH_INJECT_CO2, 0.0, 0.0, 2e-4, 1.0)
        c = w.field(ah.CO2)
        assert np.array_equal(c[0], c[1]) and np.array_equal(c[0], c[2])   # every depth alike
        assert c.max() > 1.2
        with pytest.raises(ah.AstroError):
            w.brush(ah.BRUSH_INJECT_N2, 0.0, 0.0, 1e-4, 1.0)   # N2 disabled
        with pytest.raises(ah.AstroError):
            w.fill(ah.N2, 1.0)


def test_oracle_clock_presets(oracle):
    with ah.World(oracle, capacity=10) as w:
        for preset, rates in ((ah.CLOCK_REALTIME, (1.0, 1.0)), (ah.CLOCK_MOTION, (10.0, 1.0)),
                              (ah.CLOCK_METABOLIC, (1.0, 1e4)), (ah.CLOCK_GENERATIONAL, (0.5, 1e6)),
                              (ah.CLOCK_CUSTOM, (100.0, 1.0))):
            w.set_clock_preset(preset, 500.0 if preset == ah.CLOCK_CUSTOM else 3.0, 0.5)
            assert w.time()[2:] == rates, preset   # custom (500, 0.5) clamps to (100, 1)


def test_oracle_flash_discharge_and_audit(oracle):
    """flash_step (src/sim/flash.cu:31-67): each awake charged cell loses
    min(PETROVA_FLASH_POWER * dt, E), emits nothing, runs no taxis; dormant
    cells are untouched; the audit's impulse * c / energy is 1 (the photon
    momentum identity the spin-drive scenario accepts on)."""
    kw = dict(capacity=300, chamber_w=5e-4, chamber_h=5e-4, thermal_enabled=0, emission_enabled=0,
              grid={0: (16, 16, 1), 1: (16, 16, 1), 2: (16, 16, 1), 3: (16, 16, 1)})
    with ah.World(oracle, **kw) as w:
        w.spawn(200, 5, charge_dist=ah.DIST_UNIFORM, charge_a=0.0, charge_b=1.0, awake=True)
        c = w.download()
        c["energy"][:10] = 1000.0                     # less than one tick of flash
        c["flags"][10:20] &= ~np.uint32(ah.FLAG_AWAKE)
        w.upload(c)
        e0, run0 = w.download()["energy"], w.download()["run_timer"]
        w.set_flash(True)
        w.step(1)
        c = w.download()
        de = np.minimum(3.0e6 * 1e-3, e0)            # the first ten empty within the tick
        awake = np.ones(200, bool)
        awake[10:20] = False
        assert np.array_equal(c["energy"][awake], e0[awake] - de[awake])
        assert np.all(c["energy"][:10] == 0.0)
        assert np.array_equal(c["energy"][~awake], e0[~awake])
        assert np.all(c["emit_power"][awake] == 0.0)
        assert np.array_equal(c["run_timer"], run0)   # the flash replaces taxis
        imp, dis = w.flash_audit()
        assert abs(dis - de[awake].sum()) <= 1e-6 * awake.sum()
        assert abs(imp * 299792458.0 / dis - 1.0) < 1e-9
        w.set_flash(False)
        w.step(1)
        assert w.flash_audit() == (imp, dis)          # disarmed: nothing more
