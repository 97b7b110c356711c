"""Scenarios restated in Python for tests, independently of the library's C++
scenario layer: the file is read with Python's json after stripping comments,
the reference's defaults (src/sim/scenario.cpp:139-318) are applied, and the
instantiate and drive steps (src/sim/scenario.cpp:343-468) are carried out
through a World's own entry points, so they run on either binding (the HIP
library or the CPU oracle).
"""
from __future__ import annotations

import json

import numpy as np

from conftest import ah

DT_PHYSICS = 1e-3
CANON = dict(chamber_w=0.004, chamber_h=0.004, chamber_d=6.0e-5, ambient=293.14999999999998,
             default_cells=25000, max_cells=2000000)
CLOCKS = {"realtime": ah.CLOCK_REALTIME, "motion": ah.CLOCK_MOTION, "metabolic": ah.CLOCK_METABOLIC,
          "generational": ah.CLOCK_GENERATIONAL, "custom": ah.CLOCK_CUSTOM}
BOUNDARY = {"periodic": ah.BOUNDARY_PERIODIC, "absorbing": ah.BOUNDARY_ABSORBING}
BC = {"dirichlet": 0, "neumann": 1}
PLACE = {"gaussian": ah.PLACE_GAUSSIAN, "grid": ah.PLACE_GRID, "disc": ah.PLACE_DISC}
DIST = {"uniform": ah.DIST_UNIFORM, "normal": ah.DIST_NORMAL}
STORE = {"flash": ah.STORE_FLASH, "retain": ah.STORE_RETAIN}
STIM = {"heat": ah.STIM_HEAT, "chill": ah.STIM_CHILL, "inject_co2": ah.STIM_INJECT_CO2,
        "inject_n2": ah.STIM_INJECT_N2, "set_light": ah.STIM_SET_LIGHT, "set_n2": ah.STIM_SET_N2,
        "seed_cells": ah.STIM_SEED_CELLS, "flash": ah.STIM_FLASH}
TOOLS = ["heat", "chill", "illuminate", "inject_co2", "inject_n2", "seed_cells", "seed_taumoeba", "kill",
         "charge_beam"]


def strip_comments(text: str) -> str:
    """Removes // and /* */ comments outside strings."""
    out, i, n, in_str = [], 0, len(text), False
    while i < n:
        c = text[i]
        if in_str:
            out.append(c)
            if c == "\\" and i + 1 < n:
                out.append(text[i + 1])
                i += 2
                continue
            if c == '"':
                in_str = False
            i += 1
        elif c == '"':
            in_str = True
            out.append(c)
            i += 1
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            i = n if j < 0 else j + 2
        else:
            out.append(c)
            i += 1
    return "".join(out)


def _num(obj: dict, key: str, default: float) -> float:
    v = obj.get(key, default) if isinstance(obj, dict) else default
    return float(v) if isinstance(v, (int, float)) and not isinstance(v, bool) else default


def _flag(obj: dict, key: str, default: bool) -> bool:
    v = obj.get(key, default) if isinstance(obj, dict) else default
    return v if isinstance(v, bool) else default


def _text(obj: dict, key: str, default: str) -> str:
    v = obj.get(key, default) if isinstance(obj, dict) else default
    return v if isinstance(v, str) else default


def read(text: str) -> dict:
    """The scenario's fields with the reference's defaults."""
    j = json.loads(strip_comments(text))
    ch, md, lt, ck, ob = (j.get(k) or {} for k in ("chamber", "medium", "light", "clock", "objective"))
    s = dict(
        schema=int(_num(j, "schema", 1)), id=_text(j, "id", ""), seed=int(_num(j, "seed", 20260802.0)),
        chamber_w=_num(ch, "w", CANON["chamber_w"]), chamber_h=_num(ch, "h", CANON["chamber_h"]),
        chamber_d=_num(ch, "d", CANON["chamber_d"]),
        boundary_x=BOUNDARY.get(_text(ch, "boundary_x", ""), 0), boundary_y=BOUNDARY.get(_text(ch, "boundary_y", ""), 0),
        temp_init=_num(md, "temp_init", CANON["ambient"]), co2_init=_num(md, "co2_init", 0.0),
        n2_init=_num(md, "n2_init", 0.0), ambient_temp=_num(md, "ambient_temp", CANON["ambient"]),
        thermal_bc=BC.get(_text(md, "thermal_bc", "robin"), 2),
        density_model=1 if _text(md, "density_model", "") == "water-density" else 0,
        store_disposition=STORE.get(_text(md, "store_disposition", "void"), ah.STORE_VOID),
        gravity_axis=1 if _text(md, "gravity_axis", "y") == "z" else 0,
        thermal_active=_flag(md, "thermal_active", True),
        ambient_irradiance=_num(lt, "ambient", 0.0),
        lights=[dict(dir_x=_num(l, "dir_x", 1.0), dir_y=_num(l, "dir_y", 0.0), irradiance=_num(l, "irradiance", 0.0),
                     wavelength=_num(l, "wavelength", 0.0)) for l in (lt.get("sources") or [])[:8]],
        compaction=_flag(j, "compaction", False), tau_compaction=_flag(j, "tau_compaction", False),
        clock=CLOCKS.get(_text(ck, "preset", "realtime"), ah.CLOCK_REALTIME),
        physics_rate=_num(ck, "physics_rate", 1.0), biology_rate=_num(ck, "biology_rate", 1.0),
        tools=sum(1 << TOOLS.index(t) for t in (j.get("tools") or []) if t in TOOLS),
        overrides={k: _num({"v": v}, "v", 0.0) for k, v in list((j.get("param_overrides") or {}).items())[:32]},
        run_duration_s=_num(ob, "duration_s", 0.0),
    )
    s["populations"] = []
    for p in (j.get("populations") or [])[:8]:
        cg = p.get("charge") or {}
        s["populations"].append(dict(
            kind=1 if _text(p, "kind", "") == "taumoeba" else 0, count=int(_num(p, "count", 0)),
            placement=PLACE.get(_text(p, "placement", "uniform"), ah.PLACE_UNIFORM),
            place_x=_num(p, "place_x", 0.0), place_y=_num(p, "place_y", 0.0), place_radius=_num(p, "place_radius", 0.0),
            charge_dist=DIST.get(_text(cg, "dist", "constant"), ah.DIST_CONSTANT),
            charge_a=_num(cg, "value", _num(cg, "a", 0.0)), charge_b=_num(cg, "b", 0.0),
            awake=_flag(p, "awake", False)))
    s["drive"] = []
    for e in (j.get("drive") or [])[:16]:
        v0 = _num(e, "v0", _num(e, "value", 0.0))
        s["drive"].append(dict(kind=STIM[e["kind"]], t0_s=_num(e, "t0_s", 0.0), t1_s=_num(e, "t1_s", 0.0),
                               x=_num(e, "x", 0.0), y=_num(e, "y", 0.0), radius=_num(e, "radius", 0.0),
                               v0=v0, v1=_num(e, "v1", _num(e, "value", v0))))
    s["accept"] = []
    for c in (ob.get("accept") or [])[:16]:
        s["accept"].append(dict(metric=ah.METRICS.index(c["metric"]),
                                op=ah.OPS.index(c["op"]) if c.get("op") in ah.OPS else 0,
                                value=_num(c, "value", 0.0), tol=_num(c, "tol", 0.0),
                                after_s=_num(c, "after_s", 0.0), tol_absolute=_flag(c, "tol_absolute", False)))
    return s


def capacity(s: dict) -> int:
    cells = sum(p["count"] for p in s["populations"] if p["kind"] == 0)
    return int(min(max(cells * 2 + 4096, CANON["default_cells"]), CANON["max_cells"]))


def tau_capacity(s: dict) -> int:
    tau = sum(p["count"] for p in s["populations"] if p["kind"] == 1)
    return int(min(tau * 4 + 256, 65536))


def instantiate(binding, s: dict, **base) -> ah.World:
    """scenario_instantiate through the World's entry points."""
    w = ah.World(binding, capacity=capacity(s), chamber_w=s["chamber_w"], chamber_h=s["chamber_h"],
                 chamber_d=s["chamber_d"], seed=s["seed"], co2_init=s["co2_init"], boundary_x=s["boundary_x"],
                 boundary_y=s["boundary_y"], gravity_axis=s["gravity_axis"], ambient_temp=s["ambient_temp"],
                 store_disposition=s["store_disposition"], thermal_enabled=int(s["thermal_active"]),
                 compaction_enabled=int(s["compaction"]), tau_capacity=tau_capacity(s),
                 tau_compaction_enabled=int(s["tau_compaction"]), **base)
    if s["temp_init"] != s["ambient_temp"]:
        w.fill(ah.TEMPERATURE, s["temp_init"])
    if s["n2_init"] != 0.0:
        w.fill(ah.N2, s["n2_init"])
    seed = s["seed"]
    for p in s["populations"]:
        if p["count"] <= 0:
            continue
        if p["kind"] == 1:
            w.tau_spawn(p["count"], seed)
            seed += 1
            continue
        w.spawn(p["count"], seed, placement=p["placement"], place_x=p["place_x"], place_y=p["place_y"],
                place_radius=p["place_radius"], charge_dist=p["charge_dist"], charge_a=p["charge_a"],
                charge_b=p["charge_b"], awake=p["awake"])
        seed += 1
    d = w.desc
    if s["lights"]:
        l = s["lights"][0]
        w.set_light(l["dir_x"], l["dir_y"], l["irradiance"], s["ambient_irradiance"], True)
    else:
        w.set_light(d.light_dir_x, d.light_dir_y, d.light_irradiance, s["ambient_irradiance"], bool(d.light_enabled))
    w.set_clock_preset(s["clock"], s["physics_rate"], s["biology_rate"])
    return w


def apply_drive(w: ah.World, s: dict):
    """scenario_apply_drive for the world's current tick."""
    tick, t, physics, _ = w.time()
    dt = DT_PHYSICS * physics
    d = ah.WorldDesc()
    w._check(w.b.world_desc_get(w.h, ah.C.byref(d)))
    flash = False
    for e in s["drive"]:
        open_ended = e["t1_s"] <= e["t0_s"]
        if t < e["t0_s"] or (not open_ended and t >= e["t1_s"]):
            continue
        frac = 0.0 if open_ended else (t - e["t0_s"]) / (e["t1_s"] - e["t0_s"])
        strength = e["v0"] + (e["v1"] - e["v0"]) * frac
        radius = e["radius"] if e["radius"] > 0.0 else 10.0 * (d.chamber_w + d.chamber_h)
        k = e["kind"]
        if k in (ah.STIM_HEAT, ah.STIM_CHILL, ah.STIM_INJECT_CO2, ah.STIM_INJECT_N2):
            brush = {ah.STIM_HEAT: ah.BRUSH_HEAT, ah.STIM_CHILL: ah.BRUSH_CHILL,
                     ah.STIM_INJECT_CO2: ah.BRUSH_INJECT_CO2, ah.STIM_INJECT_N2: ah.BRUSH_INJECT_N2}[k]
            w.brush(brush, e["x"], e["y"], radius, strength * dt)
        elif k == ah.STIM_SET_LIGHT:
            w.set_light(d.light_dir_x, d.light_dir_y, strength, d.ambient_irradiance, strength > 0.0)
            w._check(w.b.world_desc_get(w.h, ah.C.byref(d)))
        elif k == ah.STIM_SET_N2:
            w.fill(ah.N2, strength)
        elif k == ah.STIM_SEED_CELLS:
            deficit = int(strength) - w.counters().count
            pops = [p for p in s["populations"] if p["kind"] == 0]
            if deficit > 0 and pops:
                p = pops[0]
                w.spawn(deficit, d.seed + tick, placement=p["placement"], place_x=p["place_x"],
                        place_y=p["place_y"], place_radius=p["place_radius"], charge_dist=p["charge_dist"],
                        charge_a=p["charge_a"], charge_b=p["charge_b"], awake=p["awake"])
        else:   # the spin-drive flash, re-armed each tick while active
            flash = True
    w.set_flash(flash)


def drive_run(w: ah.World, s: dict, ticks: int):
    """Drive then step, one tick at a time."""
    for _ in range(ticks):
        apply_drive(w, s)
        w.step(1)
    return w


def brush_increment(shape, dx, dy, cx, cy, radius, amount) -> np.ndarray:
    """The smooth disc in float64 for closed-form checks (no rounding
    claims): amount * (1 - (r/R)^2)^2 inside R, per voxel centre."""
    nz, ny, nx = shape
    x = (np.arange(nx) + 0.5) * dx - 0.5 * nx * dx
    y = (np.arange(ny) + 0.5) * dy - 0.5 * ny * dy
    r = np.sqrt((x[None, :] - cx) ** 2 + (y[:, None] - cy) ** 2)
    w = np.where(r <= radius, (1.0 - (r / radius) ** 2) ** 2, 0.0)
    return np.broadcast_to(amount * w, (nz, ny, nx))
