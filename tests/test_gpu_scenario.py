"""Scenarios on the GPU.

* The stimuli a scenario applies (brush, fill, light, clock preset) against
  the CPU oracle: fields bit for bit after brushes and fills, and whole lit
  ticks after a light change within the world parity tolerances.
* The library's C++ scenario layer against the Python restatement in
  tests/scenario_py.py: a scenario instantiated by the library and one built
  step by step through the world's own entry points hold the same full state
  (state hash), and so do both after driven ticks -- the C++ runner batches
  the ticks outside stimulus windows, the restatement steps one at a time.
* The reference's own acceptance gates (T24, tools/headless.cpp:179-250) for
  all nine reference scenarios: each must ACCEPT.
"""
import numpy as np
import pytest

import scenario_py as sp
from conftest import ROOT, ah, compare_cells

pytestmark = pytest.mark.gpu

GOLDEN = ROOT / "tests" / "golden" / "scenarios"
RUNNABLE = ["bloom", "demo-herd", "first-light", "komorov", "sandbox", "shadow-garden", "spin-drive-face",
            "taumoeba", "three-percent-line"]


def text_of(name: str) -> str:
    return (GOLDEN / f"{name}.json").read_text()


@pytest.mark.parametrize("dims", ["2d", "3d"])
def test_stimuli_match_oracle(oracle, gpu, dims):
    if dims == "2d":
        kw = dict(grid={0: (128, 128, 1), 1: (64, 64, 1), 2: (32, 32, 1), 3: (128, 128, 1)}, chamber_w=1e-3,
                  chamber_h=1e-3)
    else:
        kw = dict(grid={0: (48, 40, 8), 1: (24, 20, 4), 2: (16, 16, 2), 3: (48, 40, 8)}, chamber_w=6e-4,
                  chamber_h=5e-4, chamber_d=1e-4)
    kw.update(capacity=2000, co2_init=1e-3, biology_rate=1e6, ambient_temp=300.0)
    o, g = ah.World(oracle, **kw), ah.World(gpu, **kw)
    for w in (o, g):
        w.spawn(1500, 4, charge_dist=ah.DIST_UNIFORM, charge_b=0.9, awake=True)
        w.fill(ah.N2, 0.004)
        w.brush(ah.BRUSH_HEAT, 1.1e-4, -0.7e-4, 2.6e-4, 3.75)
        w.brush(ah.BRUSH_CHILL, -2e-4, 1e-4, 1.5e-4, 0.5)
        w.brush(ah.BRUSH_INJECT_CO2, 0.0, 0.0, 80.0, 1e-4)          # a global brush
        w.brush(ah.BRUSH_INJECT_N2, 3e-4, 3e-4, 2e-4, 0.002)
    for f in (ah.TEMPERATURE, ah.CO2, ah.N2):
        assert np.array_equal(o.field(f), g.field(f)), f
    for w in (o, g):
        w.set_light(0.3, -0.9, 2.0e4, 50.0, True)
        w.set_clock_preset(ah.CLOCK_METABOLIC)
        w.step(1)
    assert o.time() == g.time()
    assert np.array_equal(o.field(ah.IRRADIANCE), g.field(ah.IRRADIANCE))
    for w in (o, g):
        w.step(2)
    a, b = o.download(), g.download()
    compare_cells(a, b)
    for name in ("x", "y", "z", "energy", "co2_held"):
        d = np.abs(a[name] - b[name]) / np.maximum(np.abs(a[name]), 1e-300)
        assert d.max() < 1e-9, name
    o.close()
    g.close()


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_flash_matches_oracle(oracle, gpu, precision):
    """The armed flash in place of taxis, with the light feeding first:
    energies bit for bit and the same audit."""
    kw = dict(capacity=3000, chamber_w=6e-4, chamber_h=6e-4, thermal_enabled=0, light_enabled=1,
              light_irradiance=1e4, precision=precision,
              grid={0: (64, 64, 1), 1: (32, 32, 1), 2: (16, 16, 1), 3: (64, 64, 1)})
    o, g = ah.World(oracle, **kw), ah.World(gpu, **kw)
    o.spawn(2500, 6, charge_dist=ah.DIST_UNIFORM, charge_b=1.0, awake=True)
    g.upload(o.download())
    for w in (o, g):
        w.set_flash(True)
        w.step(3)
        w.set_flash(False)
        w.step(2)
    a, b = o.download(), g.download()
    compare_cells(a, b)
    assert rel_max(a["energy"], b["energy"]) < 1e-12
    assert o.flash_audit() == g.flash_audit()
    o.close()
    g.close()


def rel_max(a, b):
    return float((np.abs(a - b) / np.maximum(np.abs(a), 1e-300)).max())


@pytest.mark.parametrize("name", RUNNABLE)
def test_instantiate_matches_restatement(gpu, name):
    lib_sc = ah.scenario_parse(gpu, text_of(name))
    s = sp.read(text_of(name))
    with ah.scenario_world(gpu, lib_sc) as a, sp.instantiate(gpu, s) as b:
        assert a.time() == b.time()
        da, db = a.desc, b.desc
        for k, _ in ah.WorldDesc._fields_:
            if not k.startswith("_") and k != "grid":
                assert getattr(da, k) == getattr(db, k), k
        assert a.counters().count == sum(p["count"] for p in s["populations"] if p["kind"] == 0)
        assert a.tau_counters()[0] == sum(p["count"] for p in s["populations"] if p["kind"] == 1)
        assert a.state_hash() == b.state_hash()


def windowed_scenario() -> str:
    """A small lit chamber whose script heats a disc in a window, injects CO2
    in a second, ramps the light in a third and tops up the population in an
    open-ended fourth."""
    return """{
      "id": "windows", "seed": 77,
      "chamber": {"w": 1.0e-3, "h": 1.0e-3, "d": 6.0e-5},
      "medium": {"temp_init": 300.0, "ambient_temp": 300.0, "co2_init": 1e-3, "thermal_bc": "neumann"},
      "light": {"sources": [{"dir_x": 0.0, "dir_y": 1.0, "irradiance": 5e3}], "ambient": 10.0},
      "populations": [{"count": 1200, "placement": "gaussian", "place_radius": 2e-4,
                       "charge": {"dist": "uniform", "a": 0.1, "b": 0.9}, "awake": true}],
      "compaction": true,
      "clock": {"preset": "custom", "physics_rate": 0.5, "biology_rate": 1e5},
      "drive": [
        {"kind": "heat", "t0_s": 0.02, "t1_s": 0.05, "x": 1e-4, "radius": 3e-4, "v0": 200.0, "v1": 50.0},
        {"kind": "inject_co2", "t0_s": 0.06, "t1_s": 0.07, "v0": 0.05},
        {"kind": "set_light", "t0_s": 0.03, "t1_s": 0.08, "v0": 1e3, "v1": 2e4},
        {"kind": "seed_cells", "t0_s": 0.09, "v0": 1300}
      ],
      "objective": {"accept": [{"metric": "population_live", "op": ">", "value": 0}]}
    }"""


def test_driven_ticks_match_restatement(gpu):
    """200 ticks (0.1 s at dt 5e-4) through every kind of window; the
    library runner steps the quiet stretches in batches."""
    text = windowed_scenario()
    lib_sc = ah.scenario_parse(gpu, text)
    s = sp.read(text)
    grid = dict(grid={0: (128, 128, 1), 1: (64, 64, 1), 2: (32, 32, 1), 3: (128, 128, 1)})
    with ah.scenario_world(gpu, lib_sc, **grid) as a, sp.instantiate(gpu, s, **grid) as b:
        res = ah.scenario_assert(gpu, lib_sc, a, ticks=200)
        assert res["ticks"] == 200 and res["all_pass"]
        sp.drive_run(b, s, 200)
        assert a.time() == b.time()
        ca, cb = a.counters(), b.counters()
        assert (ca.count, ca.next_id) == (cb.count, cb.next_id)
        assert cb.count >= 1300   # the top-up ran
        assert a.desc.light_irradiance == b.desc.light_irradiance
        assert a.state_hash() == b.state_hash()


def test_driven_ticks_match_oracle(gpu, oracle):
    """First-light's heat drive for 150 ticks: the GPU world driven by the
    library against the oracle driven by the restatement."""
    lib_sc = ah.scenario_parse(gpu, text_of("first-light"))
    s = sp.read(text_of("first-light"))
    grid = dict(grid={0: (128, 128, 1), 1: (64, 64, 1), 2: (32, 32, 1), 3: (128, 128, 1)})
    with ah.scenario_world(gpu, lib_sc, **grid) as g, sp.instantiate(oracle, s, **grid) as o:
        ah.scenario_assert(gpu, lib_sc, g, ticks=150)
        sp.drive_run(o, s, 150)
        assert g.time() == o.time()
        a, b = o.download(), g.download()
        compare_cells(a, b)
        for name in ("x", "y", "z", "energy"):
            d = np.abs(a[name] - b[name]) / np.maximum(np.abs(a[name]), 1e-300)
            assert d.max() < 1e-9, name
        To, Tg = o.field(ah.TEMPERATURE), g.field(ah.TEMPERATURE)
        assert np.abs(To.astype(np.float64) - Tg).max() <= 1e-4   # a few float ulps at 300 K
        assert Tg.mean() > 295.0   # the heat went in


@pytest.mark.parametrize("name", RUNNABLE)
def test_reference_acceptance(gpu, name, tmp_path):
    """The scenario's own accept block over its own horizon."""
    sc = ah.scenario_parse(gpu, text_of(name))
    with ah.scenario_world(gpu, sc) as w:
        csv = tmp_path / f"{name}.csv"
        res = ah.scenario_assert(gpu, sc, w, csv_path=csv)
        lines = [f"{c['metric']} {c['measured']:.6g} {c['op']} {c['value']:.6g} {'PASS' if c['passed'] else 'FAIL'}"
                 for c in res["checks"]]
        print(f"{name}: {res['ticks']} ticks, {res['sim_time_s']:.2f} s; " + "; ".join(lines), flush=True)
        assert res["all_pass"], lines
        if sc.accept_count:
            rows = [ln for ln in csv.read_text().splitlines() if not ln.startswith("#")]
            assert rows[0].startswith("t_s,n_live") and len(rows) == res["samples"] + 1
