"""The counter-based stream (rng_kind = ASTRO_RNG_PHILOX) in the oracle.

Pinned by:
  * the published Philox4x32-10 known-answer vectors (tests/golden/
    philox_kat.json, from the Random123 distribution);
  * the stream layout restated here over those blocks: key = seed, counter =
    (stage << 16 | block, tick, id low, id high), four words per block;
  * spawn's draw order (src/sim/cell_store.cu:42-130) replayed in Python over
    the Philox stream, exactly as for PCG32;
  * what a stream fixes in the whole step: T4 Einstein diffusion
    (tests/physics/test_motion.cpp:111) with independent axes, no per-cell
    state (rng_state stays 0, daughters included), and results that depend
    only on (seed, id, tick) -- the same cells in any slot order move alike.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from conftest import ah
from test_oracle_golden import EX, K, T_ROOM, _drag
from test_oracle_pins import SMALL

KAT = json.loads((Path(__file__).resolve().parent / "golden" / "philox_kat.json").read_text())["vectors"]


def kat_words():
    rows, outs = [], []
    for v in KAT:
        rows.append([int(h, 16) for h in v["ctr"] + v["key"]])
        outs.append([int(h, 16) for h in v["out"]])
    return np.array(rows, np.uint32), np.array(outs, np.uint32)


def test_philox_known_answers(oracle):
    words, want = kat_words()
    assert np.array_equal(oracle.philox_blocks(words), want)


def stream_restated(binding, seed, cid, tick, stage, n):
    blocks = (n + 3) // 4
    words = np.array([[(stage << 16) | b, tick & 0xFFFFFFFF, cid & 0xFFFFFFFF, cid >> 32,
                       seed & 0xFFFFFFFF, seed >> 32] for b in range(blocks)], np.uint32)
    return binding.philox_blocks(words).ravel()[:n]


@pytest.mark.parametrize("seed,cid,tick,stage", [(1, 1, 0, 1), (20260802, 77, 123456, 3),
                                                 (0xDEADBEEFCAFEF00D, (1 << 40) + 5, 2 ** 32 - 1, 2)])
def test_stream_layout(oracle, seed, cid, tick, stage):
    got = oracle.philox_stream(seed, cid, tick, stage, 23)
    assert np.array_equal(got, stream_restated(oracle, seed, cid, tick, stage, 23))
    # neighbouring counters give unrelated streams
    others = [oracle.philox_stream(seed, cid + 1, tick, stage, 23), oracle.philox_stream(seed, cid, tick + 1, stage, 23),
              oracle.philox_stream(seed, cid, tick, stage + 1, 23), oracle.philox_stream(seed + 1, cid, tick, stage, 23)]
    for o in others:
        assert np.count_nonzero(o == got) <= 1


class PhiloxDraws:
    """uniform01d / Box-Muller (src/core/rng.cuh:80-101) over a Philox stream."""

    def __init__(self, binding, seed, cid, tick, stage):
        self.words = list(stream_restated(binding, seed, cid, tick, stage, 64))

    def u01(self):
        hi = int(self.words.pop(0)) << 21
        lo = int(self.words.pop(0)) >> 11
        return float((hi ^ lo) & ((1 << 53) - 1)) * (1.0 / 9007199254740992.0)

    def uniform(self, lo, hi):
        return lo + (hi - lo) * self.u01()

    def pair(self):
        u1 = max(self.u01(), 1.0e-300)
        u2 = self.u01()
        mag = math.sqrt(-2.0 * math.log(u1))
        ang = 6.283185307179586476925286766559 * u2
        return mag * math.cos(ang), mag * math.sin(ang)


def test_spawn_draws_from_the_spawn_stream(oracle):
    seed, n = 1, 200
    with ah.World(oracle, capacity=n, grid=SMALL, rng_kind=ah.RNG_PHILOX, seed=99) as w:
        w.spawn(n, seed, charge_dist=ah.DIST_UNIFORM, charge_a=0.0, charge_b=0.2)
        got = w.download()
        d = w.desc
    a = K["CELL_RADIUS"]
    clamp = lambda v, lo, hi: lo if v < lo else (hi if v > hi else v)  # noqa: E731
    for i in range(n):
        r = PhiloxDraws(oracle, seed, i + 1, 0, ah.RNG_STAGE_SPAWN)
        x = clamp(r.uniform(-0.5 * d.chamber_w, 0.5 * d.chamber_w), -0.5 * d.chamber_w + a, 0.5 * d.chamber_w - a)
        y = clamp(r.uniform(-0.5 * d.chamber_h, 0.5 * d.chamber_h), -0.5 * d.chamber_h + a, 0.5 * d.chamber_h - a)
        z = r.uniform(-0.5 * d.chamber_d + a, 0.5 * d.chamber_d - a)
        q = clamp(r.uniform(0.0, 0.2), 0.0, 1.0)
        assert (got["x"][i], got["y"][i], got["z"][i]) == (x, y, z)
        assert got["energy"][i] == q * K["CELL_ENERGY_MAX"]
    assert not got["rng_state"].any()


def test_einstein_diffusion_with_counter_streams(oracle):
    """T4 through the whole step with Philox draws (the PCG32 version is in
    test_oracle_golden.py): 4,000 free dormant cells for 1 s; x and z carry no
    force (gravity is along y), so each has displacement variance 2 D t, and
    their displacements are uncorrelated."""
    n = 4000
    with ah.World(oracle, capacity=n, contact_enabled=0, thermal_enabled=0, emission_enabled=0, taxis_enabled=0,
                  lifecycle_enabled=0, adhesion_enabled=0, rng_kind=ah.RNG_PHILOX, seed=31, grid=SMALL) as w:
        w.spawn(n, 20260802, charge_dist=ah.DIST_CONSTANT, charge_a=0.0)
        c0 = w.download()
        w.step(int(round(1.0 / w.dt)))
        c1 = w.download()
        assert not c1["rng_state"].any()
    t_local = float(np.float32(T_ROOM))
    D = K["K_BOLTZ"] * t_local / _drag(oracle.lib, t_local)
    dx, dz = c1["x"] - c0["x"], c1["z"] - c0["z"]
    for d in (dx, dz):
        assert abs(np.mean(d * d) / (2.0 * D) - 1.0) < 0.07, np.mean(d * d) / (2.0 * D)
    msd = float(np.mean(dx * dx + dz * dz))
    assert abs(msd / (4.0 * D) - 1.0) <= EX["T4_DIFFUSIVITY_20C_TOL"]
    assert abs(np.corrcoef(dx, dz)[0, 1]) < 0.05


def test_results_depend_on_seed_id_tick_only(oracle):
    """The same cells uploaded in two slot orders, lit, awake (taxis tumbles
    and the integrator draw): every cell, found by id, ends identically, and
    a different world seed moves them differently."""
    kw = dict(capacity=800, chamber_w=1e-3, chamber_h=1e-3, light_enabled=1, light_irradiance=1e4, rng_kind=ah.RNG_PHILOX,
              seed=5, grid={0: (32, 32, 1), 1: (16, 16, 1), 2: (-1, 0, 0), 3: (32, 32, 1)}, contact_enabled=0)
    with ah.World(oracle, **kw) as w:
        w.spawn(800, 3, charge_dist=ah.DIST_UNIFORM, charge_b=0.9, awake=True)
        cells = w.download()
    # a remembered signal above any the cells now see: every cell that seeks
    # light or CO2 tumbles on the first tick
    cells["taxis_memory"][:] = 1.0e9
    perm = np.random.default_rng(0).permutation(800)
    out, tumbles = [], []
    for seed, order in ((5, np.arange(800)), (5, perm), (6, np.arange(800))):
        with ah.World(oracle, **dict(kw, seed=seed)) as w:
            w.upload({k: v[order] for k, v in cells.items()})
            w.step(1)
            tumbles.append(int(np.count_nonzero(w.download()["run_timer"] == 0.0)))
            w.step(4)
            c = w.download()
            by_id = np.argsort(c["id"])
            out.append({k: v[by_id] for k, v in c.items()})
    for k in ("x", "y", "z", "vx", "dir_x", "run_timer", "flags", "energy"):
        assert np.array_equal(out[0][k], out[1][k]), k
    assert not np.array_equal(out[0]["x"], out[2]["x"])
    assert not np.array_equal(out[0]["dir_x"], out[2]["dir_x"])
    assert tumbles[0] == tumbles[1] > 300, tumbles


def test_daughters_carry_no_stream_state(oracle):
    with ah.World(oracle, capacity=400, co2_init=0.0, grid=SMALL, rng_kind=ah.RNG_PHILOX, compaction_enabled=1) as w:
        w.spawn(200, 2)
        c = w.download()
        c["co2_held"][::2] = K["CO2_MASS_PER_DIVISION"]
        w.upload(c)
        w.stage("lifecycle", 1e-3)
        c = w.download()
        assert w.counters().count == 300
    assert not c["rng_state"].any()
