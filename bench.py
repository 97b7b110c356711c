"""Throughput of the per-step hot path on one MI355X.

Workload (the 1M-cell "plain method, scale" configuration): 1,000,000 cells
uniform in a 4 x 4 x 0.2 mm chamber; 512 x 512 x 32 FP32 temperature and CO2
grids; spatial hash with 10 um buckets; dt = 1e-4 s (physics_rate 0.1, two
temperature substeps per tick); division on with daughters offset +/-0.5 um;
compaction on. To make division and compaction do work every tick:
  * 10 % of the cells are awake with a near-empty store, so the thermostat
    starves a trickle of them each tick (about 100 deaths per tick);
  * CO2 starts at 1 kg/m^3, the biology clock runs at 1e5, and each cell
    starts with a uniform fraction of the division quota, so a trickle divide;
  * the other 90 % are dormant with log-uniform charge in [1e-3, 0.2].
The light is off. Positions and velocities are stored FP32 (the integrator's
drag, coefficients and Gaussians in FP32, displacement sums in FP64),
energies and CO2 stores FP64, fields FP32.

`value` = cells x ticks / second over the timed ticks (the cell count of each
tick summed on the device), with all state resident in HBM; the ticks are
replayed from a captured hipGraph. Afterwards a shorter pass re-runs the same
ticks eagerly with HIP events around every kernel on the world's stream to
give per-kernel times and achieved bandwidth against algorithmic bytes.
Each kernel's achieved bandwidth is taken on the frozen per-stage byte basis
(roofline.STAGE_BYTES); `roofline` is the kernel that takes the most time per
tick, `step_roofline` the whole tick. The CPU baseline is the CPU restatement
of the same step (FP64, the reference's stage order) on the same workload for
a bounded number of ticks, single-threaded and on every core with OpenMP.
Per-cell streams are Philox4x32-10 by default (--rng pcg32 for the
reference's generator).

`configs` holds the same measurements for the small (2,000 cells, 64^3) and
real-shape (25,000 awake lit cells, 256 x 256 x 64) plain workloads.

The `ensemble` object measures the batched-ensemble workload: 256 independent
chambers (1 x 1 x 0.2 mm) with cell counts log-uniform in [500, 50,000] drawn
from seed 7, each with its own 64^3 FP32 temperature and irradiance grids and a
random in-plane light direction, every cell awake. All chambers advance
together, each stage one batched launch over every chamber and each step one
graph launch. The steps are timed before and after the autotuner picks the
launch shapes; the same chambers stepped one after another with their own
graphs give the sequential figure beside it.

The `surrogate` object measures the learned field surrogate at 128^3: it is
trained on 2,000 plain trajectories (about 20 s), then timed per call in its
three modes (plain FTCS, net, gated) and compared on held-out states with the
FP64 CPU plain path, on the full fields and on the increments.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
PKG = ROOT / "astrophage-microscope_amd"
sys.path.insert(0, str(PKG))

import numpy as np  # noqa: E402

import astro_hot as ah  # noqa: E402
import roofline  # noqa: E402

C_INT = ah.C.c_int32

METRIC = "cell-steps/s at 1M cells; field steps/s surrogate vs FTCS at 1e-3 rel err"
HBM_PEAK_GBS = 8000.0
QUOTA = 2.1e-14


def workload(binding: ah.Binding, cells: int, precision: int, rng_kind: int = 0):
    return binding.desc(
        capacity=int(cells * 1.25), seed=1, chamber_w=4e-3, chamber_h=4e-3, chamber_d=2e-4,
        precision=precision, hash_cell_size=1e-5, rng_kind=rng_kind,
        grid={ah.TEMPERATURE: (512, 512, 32), ah.CO2: (512, 512, 32), ah.N2: (-1, 0, 0),
              ah.IRRADIANCE: (64, 64, 1)},
        physics_rate=0.1, biology_rate=1e5, co2_init=1.0, division_offset=0.5e-6, compaction_enabled=1,
    )


def populate(w: ah.World, cells: int):
    awake = cells // 10
    w.spawn(cells - awake, 1, charge_dist=ah.DIST_LOG_UNIFORM, charge_a=1e-3, charge_b=0.2)
    w.spawn(awake, 2, charge_dist=ah.DIST_UNIFORM, charge_a=0.0, charge_b=1e-10, awake=True)
    c = w.download()
    c["co2_held"] = np.random.default_rng(1).uniform(0.0, 1.0, cells) * QUOTA
    w.upload(c)
    ctr = w.counters()
    ctr.next_id = cells + 1
    w.set_counters(ctr)


def workload_c1(binding: ah.Binding, precision: int, rng_kind: int = 0, cells: int = 2000):
    """Plain method, small: 2,000 cells in a 550 x 550 x 100 um slab, 64^3 T
    at 369.0 K (just below the 369.565 K ignition point), no light,
    dt = 1e-4 s, 22 um hash."""
    return binding.desc(
        capacity=cells, seed=1, chamber_w=550e-6, chamber_h=550e-6, chamber_d=100e-6, precision=precision,
        rng_kind=rng_kind, ambient_temp=369.0, physics_rate=0.1,
        grid={ah.TEMPERATURE: (64, 64, 64), ah.CO2: (64, 64, 64), ah.N2: (-1, 0, 0), ah.IRRADIANCE: (64, 64, 64)})


def populate_c1(w: ah.World, cells: int = 2000):
    w.spawn(cells, 1, charge_dist=ah.DIST_UNIFORM, charge_a=0.0, charge_b=0.2)


C2_GRID = (256, 256, 64)
C2_HOT_SIGMA = 40e-6
C2_HOT_DT = 5.0


def workload_c2(binding: ah.Binding, precision: int, rng_kind: int = 0, cells: int = 25_000):
    """Plain method, real shape: 25,000 awake cells in a 550 x 550 x 200 um
    box, 256 x 256 x 64 FP32 temperature and nutrient (CO2) grids, collimated
    light along +x at 1e4 W/m^2 (the reference's lit test value,
    tests/physics/test_emission.cu:84) swept over a 256 x 256 x 64 grid,
    dt = 1e-4 s, 22 um hash."""
    return binding.desc(
        capacity=int(cells * 1.25), seed=1, chamber_w=550e-6, chamber_h=550e-6, chamber_d=200e-6,
        precision=precision, rng_kind=rng_kind, physics_rate=0.1,
        grid={ah.TEMPERATURE: C2_GRID, ah.CO2: C2_GRID, ah.N2: (-1, 0, 0), ah.IRRADIANCE: C2_GRID},
        light_enabled=1, light_irradiance=1e4, light_dir_x=1.0, light_dir_y=0.0)


def c2_hot_spot(w: ah.World) -> np.ndarray:
    """Ambient plus a 5 K Gaussian (sigma 40 um) at the box centre, on the
    temperature grid's cell centres, z-major (z, y, x)."""
    (nx, ny, nz), (dx, dy, dz) = w.field_dims(ah.TEMPERATURE)
    d = w.desc
    x = (np.arange(nx) + 0.5) * dx - 0.5 * d.chamber_w
    y = (np.arange(ny) + 0.5) * dy - 0.5 * d.chamber_h
    z = (np.arange(nz) + 0.5) * dz - 0.5 * d.chamber_d
    r2 = z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2
    return (d.ambient_temp + C2_HOT_DT * np.exp(-r2 / (2.0 * C2_HOT_SIGMA ** 2))).astype(np.float32).ravel()


def populate_c2(w: ah.World, cells: int = 25_000):
    w.spawn(cells, 1, charge_dist=ah.DIST_LOG_UNIFORM, charge_a=1e-3, charge_b=0.2, awake=True)
    w.set_field(ah.TEMPERATURE, c2_hot_spot(w))


def population(w: ah.World) -> dict:
    s = w.stats()
    return {"n": s["n_live"] + s["n_dead"], "alive": s["n_live"], "awake": s["n_awake"],
            "dead": s["n_dead"], "divisions": s["divisions_this_window"], "deaths": s["deaths_this_window"]}


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_binding(all_cores: bool) -> ah.Binding:
    name = "liboracle_omp.so" if all_cores else "liboracle.so"
    lib = ROOT / "oracle" / name
    if not lib.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle"), "omp" if all_cores else "oracle"], check=True,
                       stdout=subprocess.DEVNULL)
    return ah.Binding(lib, "oracle_", {"threads": (C_INT, [])})


def cpu_baseline(make, fill, serial_ticks: int, omp_ticks: int, what: str) -> dict:
    """The oracle (FP64, the reference's stage order) on the same workload,
    timed on this host: single-threaded, and built with OpenMP on every core
    it is given (bit-identical results). One warm-up tick each, then a bounded
    sample of ticks; value = the all-cores rate."""
    out = {}
    for all_cores, ticks in ((False, serial_ticks), (True, omp_ticks)):
        b = oracle_binding(all_cores)
        with make(b) as w:
            fill(w)
            w.step(1)
            c0 = w.counters().cell_steps
            t0 = time.perf_counter()
            w.step(ticks)
            dt = time.perf_counter() - t0
            done = w.counters().cell_steps - c0
        out[all_cores] = (done / dt, int(b.threads()), ticks, dt)
    rate, threads, ticks, dt = out[True]
    rate1, _, ticks1, dt1 = out[False]
    return {"value": rate, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": f"{what}: {ticks} ticks after 1 warm-up tick on {threads} threads ({dt:.1f} s), "
                      f"FP64 CPU restatement built with OpenMP",
            "single_thread": {"value": rate1, "cores": 1, "sample": f"{ticks1} ticks ({dt1:.1f} s)"},
            "cpu_model": cpu_model(), "hardware_threads": os.cpu_count()}


def world_counts(w: ah.World, before: dict, after: dict, ticks: int, precision: int, philox: bool) -> dict:
    """Population and grid figures of a measured interval, for the byte tables."""
    nx, ny, nz, _ = w.hash_dims()
    pts = {}
    for f in (ah.TEMPERATURE, ah.CO2, ah.N2, ah.IRRADIANCE):
        (gx, gy, gz), _ = w.field_dims(f)
        if gx > 0:
            pts[f] = gx * gy * gz
    return {"n": 0.5 * (before["n"] + after["n"]), "alive": 0.5 * (before["alive"] + after["alive"]),
            "awake": 0.5 * (before["awake"] + after["awake"]), "dead": after["deaths"] / max(ticks, 1),
            "births": after["divisions"] / max(ticks, 1), "compacting": 1.0 if after["deaths"] > 0 else 0.0,
            "p": 4 if precision == ah.FP32_HOT else 8, "buckets": nx * ny * nz, "pts": pts,
            "sampled_fields": 2, "resort_frac": 1.0 / 16.0, "philox": philox}


def measure_world(w: ah.World, steps: int, warmup: int, profile_steps: int, precision: int, philox: bool,
                  traffic: dict | None = None) -> dict:
    """Graph-replayed ticks timed end to end, then the same ticks eager with
    HIP events around every kernel on the world's stream. Each kernel's
    achieved bandwidth is its bytes on the frozen survey basis
    (roofline.survey_kernel_bytes) over its average launch time; the
    builder's own count is printed beside it."""
    w.step(warmup)
    w.sync()
    c0 = w.counters()
    t0 = time.perf_counter()
    w.step(steps)
    w.sync()
    elapsed = time.perf_counter() - t0
    c1 = w.counters()
    value = (c1.cell_steps - c0.cell_steps) / elapsed

    before = population(w)
    w.profile(True)
    w.profile_reset()
    w.step(profile_steps)
    kern = w.profile_read()
    w.profile(False)
    after = population(w)
    counts = world_counts(w, before, after, profile_steps, precision, philox)
    stages, total_ms = {}, 0.0
    for name, (launches, ms) in kern.items():
        total_ms += ms
        per = ms / max(launches, 1)
        basis = roofline.survey_kernel_bytes(name, counts)
        own = roofline.kernel_bytes(name, counts)
        stages[name] = {"launches_per_tick": launches / profile_steps, "avg_us": round(per * 1e3, 2),
                        "bytes_per_launch": basis,
                        "achieved_GBs": round(basis / (per * 1e-3) / 1e9, 1) if per > 0 else None,
                        "builder_bytes_per_launch": own,
                        "builder_GBs": round(own / (per * 1e-3) / 1e9, 1) if per > 0 else None}
    dom = max(kern.items(), key=lambda kv: kv[1][1])[0]
    d = stages[dom]
    roof = {"kernel": dom, "bound": "hbm", "achieved": d["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(d["achieved_GBs"] / HBM_PEAK_GBS, 4) if d["achieved_GBs"] else None,
            "traffic": (traffic or {}).get(dom), "avg_launch_us": d["avg_us"],
            "algo_bytes_per_launch": d["bytes_per_launch"], "basis": "frozen per-stage bytes (roofline.STAGE_BYTES)",
            "builder_achieved": d["builder_GBs"], "builder_bytes_per_launch": d["builder_bytes_per_launch"]}
    per_tick = sum(roofline.survey_kernel_bytes(k, counts) * v[0] for k, v in kern.items()) / profile_steps
    own_tick = sum(roofline.kernel_bytes(k, counts) * v[0] for k, v in kern.items()) / profile_steps
    ms_step = elapsed / steps * 1e3
    step_roof = {"bytes_per_tick": per_tick, "achieved_GBs": round(per_tick / (ms_step * 1e-3) / 1e9, 1),
                 "frac": round(per_tick / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                 "builder_bytes_per_tick": own_tick, "eager_ms_per_tick": round(total_ms / profile_steps, 4)}
    return {"value": value, "ms_per_step": ms_step, "steps": steps, "warmup": warmup, "c0": c0, "c1": c1,
            "roofline": roof, "step_roofline": step_roof, "stages": stages}


ENSEMBLE_GRID = {ah.TEMPERATURE: (64, 64, 64), ah.CO2: (16, 16, 4), ah.N2: (-1, 0, 0), ah.IRRADIANCE: (64, 64, 64)}


def ensemble_chambers(binding: ah.Binding, chambers: int, seed: int = 7) -> list[ah.World]:
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(500.0), np.log(50_000.0), chambers)).astype(np.int64)
    angles = rng.uniform(0.0, 2.0 * np.pi, chambers)
    worlds = []
    for i, (n, ang) in enumerate(zip(counts, angles)):
        d = binding.desc(capacity=int(n * 1.25) + 64, seed=seed * 1000 + i, chamber_w=1e-3, chamber_h=1e-3,
                         chamber_d=2e-4, precision=ah.FP32_HOT, grid=ENSEMBLE_GRID, physics_rate=0.1,
                         light_enabled=1, light_irradiance=1e4, light_dir_x=float(np.cos(ang)),
                         light_dir_y=float(np.sin(ang)))
        w = ah.World(binding, d)
        w.spawn(int(n), seed * 1000 + i, charge_dist=ah.DIST_LOG_UNIFORM, charge_a=1e-3, charge_b=0.2, awake=True)
        worlds.append(w)
    return worlds


def total_cell_steps(worlds: list[ah.World]) -> int:
    return sum(w.counters().cell_steps for w in worlds)


def ensemble_cpu_baseline(chambers: int, sample: int, ticks: int) -> dict:
    """The oracle on every core, stepping the first `sample` chambers of the
    ensemble one after another for `ticks` ticks each."""
    b = oracle_binding(True)
    worlds = ensemble_chambers(b, chambers)
    done, secs = 0, 0.0
    for w in worlds[:sample]:
        w.step(1)
        c0 = w.counters().cell_steps
        t0 = time.perf_counter()
        w.step(ticks)
        secs += time.perf_counter() - t0
        done += w.counters().cell_steps - c0
    for w in worlds:
        w.close()
    return {"value": done / secs, "unit": "cell-steps/s", "cores": int(b.threads()), "kind": "port",
            "sample": f"the first {sample} of the {chambers} chambers, {ticks} ticks each after 1 warm-up tick "
                      f"({secs:.1f} s), FP64 CPU restatement built with OpenMP",
            "cpu_model": cpu_model(), "hardware_threads": os.cpu_count()}


def bench_ensemble(binding: ah.Binding, chambers: int, steps: int, warmup: int, profile_steps: int,
                   seq_steps: int, sequential: bool = True, cpu: bool = True) -> dict:
    worlds = ensemble_chambers(binding, chambers)
    cells = [w.counters().count for w in worlds]
    e = ah.Ensemble(binding, worlds)

    def timed(n: int):
        e.step(warmup)
        e.sync()
        c = total_cell_steps(worlds)
        t = time.perf_counter()
        e.step(n)
        e.sync()
        return time.perf_counter() - t, total_cell_steps(worlds) - c

    untuned_elapsed, untuned_done = timed(steps)
    tuning = e.autotune(2)
    elapsed, done = timed(steps)

    # The same chambers one after another, each replaying its own graph.
    seq_elapsed, seq_done = float("nan"), 0
    if sequential:
        for w in worlds:
            w.step(2)
            w.sync()
        s0 = total_cell_steps(worlds)
        t1 = time.perf_counter()
        for w in worlds:
            w.step(seq_steps)
            w.sync()
        seq_elapsed = time.perf_counter() - t1
        seq_done = total_cell_steps(worlds) - s0

    # Per-stage pass: eager batched launches with HIP events on the ensemble's stream.
    e.profile(True)
    e.profile_reset()
    e.step(profile_steps)
    kern = e.profile_read()
    e.profile(False)
    per_chamber = []
    for w in worlds:
        st = w.stats()
        per_chamber.append({"n": st["n_live"] + st["n_dead"], "alive": st["n_live"], "awake": st["n_awake"],
                            "dead": 0, "births": 0, "compacting": 0.0, "p": 4, "buckets": int(np.prod(w.hash_dims()[:3])),
                            "pts": {0: 64 ** 3, 1: 16 * 16 * 4, 3: 64 ** 3}, "sampled_fields": 2,
                            "resort_frac": 1.0 / 16.0, "philox": False})
    # Per tick: a stage may be several batched launches (substeps, or one
    # launch per kernel variant when chambers differ, as the two sweep
    # directions do); its algorithmic bytes are every chamber's, times the
    # chamber's substeps.
    substepped = {"thermal", "deposit_apply", "ftcs_T", "ftcs_CO2", "ftcs_N2"}
    stages = {}
    for name, (launches, ms) in kern.items():
        mult = launches / profile_steps if name in substepped else 1.0
        algo = mult * sum(roofline.survey_kernel_bytes(name, c) for c in per_chamber)
        own = mult * sum(roofline.kernel_bytes(name, c) for c in per_chamber)
        per_tick = ms / profile_steps
        stages[name] = {"launches_per_tick": launches / profile_steps, "us_per_tick": round(per_tick * 1e3, 2),
                        "algo_bytes_per_tick": algo,
                        "achieved_GBs": round(algo / (per_tick * 1e-3) / 1e9, 1) if per_tick > 0 else None,
                        "builder_bytes_per_tick": own}
    algo_step = sum(v["algo_bytes_per_tick"] for v in stages.values())
    dom = max(kern.items(), key=lambda kv: kv[1][1])[0]
    value = done / elapsed
    seq_value = seq_done / seq_elapsed if sequential else None
    out = {
        "workload": "batched ensemble: 256 chambers 1x1x0.2 mm, cells log-uniform [500, 50000] (seed 7), "
                    "64^3 FP32 T + irradiance each, random in-plane light, all awake, dt 1e-4 s",
        "chambers": len(worlds), "cells": int(sum(cells)), "cells_min": int(min(cells)), "cells_max": int(max(cells)),
        "metric": "cell-steps/s (all chambers)", "value": value, "steps": steps, "warmup": warmup,
        "ms_per_step": elapsed / steps * 1e3, "chamber_steps_per_s": len(worlds) * steps / elapsed,
        "untuned": {"value": untuned_done / untuned_elapsed, "ms_per_step": untuned_elapsed / steps * 1e3,
                    "how": "2048 blocks per grid-stride stage, widest FTCS row group"},
        "tuning": tuning,
        "sequential": {"value": seq_value, "ms_per_step_all_chambers": seq_elapsed / seq_steps * 1e3,
                       "steps": seq_steps, "how": "each chamber replays its own graph, one chamber after another"},
        "speedup_vs_sequential": value / seq_value if seq_value else None,
        "step_roofline": {"algo_bytes_per_step": algo_step,
                          "achieved_GBs_graph": round(algo_step / (elapsed / steps) / 1e9, 1),
                          "frac_graph": round(algo_step / (elapsed / steps) / 1e9 / HBM_PEAK_GBS, 4),
                          "dominant_stage": dom},
        "stages": stages,
        "cpu_baseline": ensemble_cpu_baseline(chambers, 8, 2) if cpu else None,
    }
    e.close()
    for w in worlds:
        w.close()
    return out


SURROGATE_FLOPS_PER_VOXEL = 2 * 27 * (8 * 32 + 32 * 32 + 32 * 32 + 32 * 2)   # one call, useful MACs x 2


def surrogate_cpu_reference(states, params: dict) -> dict:
    """CPU leg of the surrogate measurement: the FP64 plain path (the oracle)
    on the held-out states, timed, and used as the reference for the error
    figures."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import surrogate_oracle as so
    out, secs, steps = [], 0.0, 0
    for T0, C0, S, occ in states:
        t0 = time.perf_counter()
        Tw, Cw = so.plain_advance(T0, C0, S, occ, params)
        secs += time.perf_counter() - t0
        steps += params["steps_per_call"]
        out.append((Tw, Cw))
    return {"refs": out, "field_steps_per_s": steps / secs, "rel_l2": so.rel_l2}


def bench_surrogate(binding: ah.Binding, trajectories: int, epochs: int, calls: int) -> dict:
    s = ah.Surrogate(binding)
    t0 = time.perf_counter()
    loss = s.train(trajectories, epochs=epochs, lr=1e-3, seed=1)
    train_s = time.perf_counter() - t0

    # held-out states: the net's and the FP32 plain path's results
    states, net, plain = [], [], []
    for seed in (100001, 100002, 100003):
        s.sample(seed)
        st = s.state()
        states.append(st)
        s.advance(ah.SURROGATE_NET)
        net.append(s.state()[:2])
        s.set_state(*st)
        s.advance(ah.SURROGATE_PLAIN)
        plain.append(s.state()[:2])
    cpu = surrogate_cpu_reference(states, s.params())
    rel = cpu["rel_l2"]

    def errs(res):
        e = {"temp_field": [], "temp_increment": [], "nutrient_field": [], "nutrient_increment": []}
        for (T1, C1), (T0, C0, _, _), (Tw, Cw) in zip(res, states, cpu["refs"]):
            e["temp_field"].append(rel(T1, Tw))
            e["temp_increment"].append(rel(T1.astype(np.float64) - T0, Tw - T0))
            e["nutrient_field"].append(rel(C1, Cw))
            e["nutrient_increment"].append(rel(C1.astype(np.float64) - C0, Cw - C0))
        return {k: float(np.mean(v)) for k, v in e.items()}

    timing = {}
    s.sample(7)
    for name, mode in (("plain", ah.SURROGATE_PLAIN), ("net", ah.SURROGATE_NET), ("gated", ah.SURROGATE_GATED)):
        s.advance(mode)
        s.state()
        t0 = time.perf_counter()
        for _ in range(calls):
            s.advance(mode)
        s.state()
        timing[name] = (time.perf_counter() - t0) / calls
    c = s.counters()
    steps = s.desc.steps_per_call
    n = s.n
    net_tflops = SURROGATE_FLOPS_PER_VOXEL * n / timing["net"] / 1e12
    # the plain path's compulsory bytes: per substep and voxel T, C, S and
    # occupancy in, T' and C' out, 4 bytes each
    plain_bytes = 24.0 * steps * n
    net_at_peak = SURROGATE_FLOPS_PER_VOXEL * n / 2.5e15
    plain_at_peak = plain_bytes / (HBM_PEAK_GBS * 1e9)
    out = {
        "workload": "learned field surrogate: 128^3 T + nutrient, 550x550x200 um, 16 FTCS substeps of 6.25 us per "
                    "call, states with Gaussian hot/cool spots and a clustered 25k-cell culture",
        "metric": "field steps/s (16 substeps of both fields per call)",
        "plain_ms_per_call": timing["plain"] * 1e3, "net_ms_per_call": timing["net"] * 1e3,
        "gated_ms_per_call": timing["gated"] * 1e3,
        "field_steps_per_s": {"plain": steps / timing["plain"], "net": steps / timing["net"],
                              "gated": steps / timing["gated"]},
        "net_vs_plain": timing["plain"] / timing["net"],
        "error_vs_fp64": {"net": errs(net), "plain_fp32": errs(plain),
                          "held_out_states": 3},
        "gate": {"judge_every": s.desc.judge_every, "judge_tol": s.desc.judge_tol,
                 "calls": c["calls"], "fallbacks": c["fallbacks"], "last_judge_err": c["last_judge_err"]},
        "training": {"trajectories": trajectories, "epochs": epochs, "seconds": train_s, "final_loss": loss},
        "roofline": {"kernel": "conv stack (whole net call)", "bound": "mfma",
                     "achieved": round(net_tflops, 1), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(net_tflops / 2500.0, 4),
                     "flops_per_call": SURROGATE_FLOPS_PER_VOXEL * n},
        "plain_roofline": {"kernel": "k_plain_substep (whole plain call)", "bound": "hbm",
                           "achieved": round(plain_bytes / timing["plain"] / 1e9, 1), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(plain_bytes / timing["plain"] / 1e9 / HBM_PEAK_GBS, 4),
                           "bytes_per_call": plain_bytes},
        "ceiling": {"net_us_at_mfma_peak": round(net_at_peak * 1e6, 1),
                    "plain_us_at_hbm_peak": round(plain_at_peak * 1e6, 1),
                    "net_vs_plain_at_both_peaks": round(plain_at_peak / net_at_peak, 3)},
        "cpu_baseline": {"value": cpu["field_steps_per_s"], "unit": "field steps/s", "cores": 1, "kind": "port",
                         "sample": "FP64 numpy plain path on the 3 held-out states (16 substeps each)"},
    }
    s.close()
    return out


RNG = {"philox": ah.RNG_PHILOX, "pcg32": ah.RNG_PCG32}


def bench_plain(gpu: ah.Binding, name: str, make, fill, precision: int, rng: str, steps: int, warmup: int,
                profile_steps: int, cpu: tuple | None, traffic: dict | None = None) -> dict:
    """One plain-method configuration: GPU throughput and rooflines, and the
    CPU baseline on a bounded sample of the same workload."""
    w = make(gpu)
    fill(w)
    m = measure_world(w, steps, warmup, profile_steps, precision, rng == "philox", traffic)
    w.close()
    c0, c1 = m.pop("c0"), m.pop("c1")
    m["divisions"] = c1.divisions_total - c0.divisions_total
    m["deaths"] = c1.deaths_total - c0.deaths_total
    m["cells_end"] = c1.count
    if cpu is not None:
        serial_ticks, omp_ticks = cpu
        m["cpu_baseline"] = cpu_baseline(make, fill, serial_ticks, omp_ticks, name)
    return m


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=int, default=1_000_000)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--rng", choices=sorted(RNG), default="philox",
                    help="per-cell streams: counter-based Philox4x32-10 (default) or the reference's PCG32")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-ticks", type=int, nargs=2, default=[4, 12], metavar=("SERIAL", "ALL_CORES"),
                    help="ticks of the 1M-cell CPU baseline, single-threaded and on all cores")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the small and real-shape lines")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-only", action="store_true")
    ap.add_argument("--ensemble-chambers", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--no-sequential", action="store_true", help="skip the per-world sequential ensemble figure")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--surrogate-only", action="store_true")
    ap.add_argument("--surrogate-trajectories", type=int, default=2000)
    a = ap.parse_args()

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(one chamber per GPU; see the sharding notes in DESIGN.md).", file=sys.stderr)
        return 2

    gpu = ah.load()
    if a.ensemble_only:
        print(json.dumps(bench_ensemble(gpu, a.ensemble_chambers, a.ensemble_steps, 10, 10, 10,
                                        sequential=not a.no_sequential, cpu=not a.no_cpu)))
        return 0
    if a.surrogate_only:
        print(json.dumps(bench_surrogate(gpu, a.surrogate_trajectories, 1, 50)))
        return 0
    precision = ah.FP32_HOT if a.precision == "fp32" else ah.FP64_CONTRACT
    rng = RNG[a.rng]
    traffic = None
    tfile = ROOT / "profiles" / "traffic_c3.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text())

    # the 1M-cell scale workload: the headline line
    big = bench_plain(gpu, f"{a.cells}-cell scale workload", lambda b: ah.World(b, workload(b, a.cells, precision, rng)),
                      lambda w: populate(w, a.cells), precision, a.rng, a.steps, a.warmup, a.profile_steps,
                      None if a.no_cpu else tuple(a.cpu_ticks), traffic)
    print(json.dumps({"progress": "scale workload", "value": big["value"], "ms_per_step": big["ms_per_step"]}),
          file=sys.stderr, flush=True)

    configs = {}
    if not a.no_configs:
        configs["small"] = bench_plain(
            gpu, "2,000-cell small workload", lambda b: ah.World(b, workload_c1(b, precision, rng)), populate_c1,
            precision, a.rng, 1000, 20, 20, None if a.no_cpu else (100, 300))
        configs["small"]["workload"] = ("plain method, small: 2,000 cells in 550x550x100 um, 64^3 FP32 T at 369.0 K, "
                                        "no light, dt 1e-4 s, 1,000 steps, seed 1")
        configs["real_shape"] = bench_plain(
            gpu, "25,000-cell real-shape workload", lambda b: ah.World(b, workload_c2(b, precision, rng)), populate_c2,
            precision, a.rng, 2000, 20, 10, None if a.no_cpu else (2, 6))
        configs["real_shape"]["workload"] = (
            "plain method, real shape: 25,000 awake cells in 550x550x200 um, log-uniform charge [1e-3, 0.2], "
            "256x256x64 FP32 T + nutrient with a 5 K sigma-40-um hot spot, light along +x at 1e4 W/m^2, "
            "dt 1e-4 s, 2,000 steps")
        for k, v in configs.items():
            print(json.dumps({"progress": k, "value": v["value"], "ms_per_step": v["ms_per_step"]}), file=sys.stderr,
                  flush=True)

    ens = None
    if not a.no_ensemble:
        ens = bench_ensemble(gpu, a.ensemble_chambers, a.ensemble_steps, 10, 10, 10, cpu=not a.no_cpu)
        print(json.dumps({"progress": "ensemble", "value": ens["value"]}), file=sys.stderr, flush=True)

    sur = None
    if not a.no_surrogate:
        sur = bench_surrogate(gpu, a.surrogate_trajectories, 1, 50)

    out = {
        "metric": METRIC, "value": big["value"], "unit": "cell-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": big["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if precision == ah.FP32_HOT else "f64", "data": "synthetic",
        "config": {"workload": "plain method, scale: 1M cells, 4x4x0.2 mm, 512x512x32 FP32 T+CO2, 10 um hash, "
                               "division + compaction on, dt 1e-4 s",
                   "cells": a.cells, "cells_end": big["cells_end"], "storage": a.precision, "rng": a.rng,
                   "arithmetic": "fp32 integrator transcendentals, fp64 displacement sums and energies, fp32 fields",
                   "graph_replay": True, "divisions": big["divisions"], "deaths": big["deaths"]},
        "roofline": big["roofline"],
        "step_roofline": big["step_roofline"],
        "cpu_baseline": big.get("cpu_baseline"),
        "stages": big["stages"],
        "configs": configs,
        "ensemble": ens,
        "surrogate": sur,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
