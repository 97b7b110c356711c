"""Predation on the GPU (src/sim/predation.cu) against the CPU oracle
(oracle/astro_oracle.cpp, itself pinned in test_predation_oracle.py and
test_oracle_golden.py).

Bars, in the FP64 layout:
  * spawn: ids, flags, streams, positions and every scalar bit for bit;
    headings (normalized Box-Muller draws through libm) within a float ulp;
  * whole ticks of a crowded chamber with predators hunting, digesting,
    dying of nitrogen, dividing and compacting: every integer (cell flags
    and death causes, predator ids, flags, generations, targets, streams,
    counts) bit for bit; positions, velocities and biomass within 1e-9
    relative (libm ulps in the OU coefficients and Gaussians);
  * stats (predator count, mean tolerance) equal;
  * snapshots carry the predators: a restored world continues bit for bit.
In the FP32 layout the run must be reproducible and eat comparably.
"""
import numpy as np
import pytest

import asph
from conftest import ah, compare_cells

pytestmark = pytest.mark.gpu

SMALL = dict(chamber_w=3e-4, chamber_h=3e-4, chamber_d=6e-5, emission_enabled=0, taxis_enabled=0,
             thermal_enabled=0, grid={0: (32, 32, 1), 1: (16, 16, 1), 2: (16, 16, 1), 3: (32, 32, 1)})


def pair(oracle, gpu, **kw):
    kw = dict(SMALL, **kw)
    return ah.World(oracle, **kw), ah.World(gpu, **kw)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float((np.abs(a - b) / np.maximum(np.abs(a), 1e-300)).max()) if a.size else 0.0


def compare_tau(a, b, tol=1e-9):
    for k in ("id", "flags", "generation", "target", "rng_state"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("x", "y", "z", "vx", "vy", "vz", "biomass", "prey_biomass"):
        assert rel(a[k], b[k]) < tol, (k, rel(a[k], b[k]))
    for k in ("digest_timer", "tolerance", "density_ema", "run_timer", "dir_x", "dir_y", "dir_z"):
        d = np.abs(a[k].astype(np.float64) - b[k])
        assert d.max(initial=0.0) <= 2e-6 * np.abs(a[k]).max(initial=1.0), (k, d.max(initial=0.0))


def test_spawn_matches_oracle(oracle, gpu):
    o, g = pair(oracle, gpu, capacity=10, tau_capacity=5000)
    for w in (o, g):
        w.tau_spawn(3000, 77)
        w.tau_spawn(1000, 78)
    a, b = o.tau_download(), g.tau_download()
    assert o.tau_counters() == g.tau_counters() == (4000, 5000, 4001)
    for k in ("id", "flags", "rng_state", "x", "y", "z", "biomass", "tolerance", "generation", "target"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("dir_x", "dir_y", "dir_z"):
        assert np.abs(a[k] - b[k]).max() <= 1.2e-7, k
    with pytest.raises(ah.AstroError) as e:
        g.tau_spawn(1001, 3)
    assert e.value.status == ah.CAPACITY_EXCEEDED and g.tau_counters()[0] == 4000
    o.close()
    g.close()


def crowded(binding, precision, seed=11, cells=2000, tau=40, **kw):
    w = ah.World(binding, **dict(SMALL, capacity=2600, tau_capacity=128, co2_init=0.0, biology_rate=1e5,
                                 seed=seed, precision=precision, tau_compaction_enabled=1, **kw))
    w.spawn(cells, seed, charge_dist=ah.DIST_CONSTANT, charge_a=0.03)
    w.tau_spawn(tau, seed)
    return w


def seeded_pair(oracle, gpu, **kw):
    """The oracle's spawn uploaded into both, so the worlds start identical;
    some predators start well fed (they divide), N2 is lethal to the least
    tolerant."""
    o = crowded(oracle, ah.FP64_CONTRACT, **kw)
    g = ah.World(gpu, **dict(SMALL, capacity=2600, tau_capacity=128, co2_init=0.0, biology_rate=1e5, seed=11,
                             tau_compaction_enabled=1, **kw))
    cells = o.download()
    t = o.tau_download()
    rng = np.random.default_rng(5)
    t["biomass"][rng.random(len(t["id"])) < 0.3] = 1.99e-13   # one meal from dividing
    t["tolerance"][:] = rng.uniform(0.0, 0.6, len(t["id"])).astype(np.float32)
    o.tau_upload(t)
    g.upload(cells)
    g.tau_upload(t)
    for w in (o, g):
        w.fill(ah.N2, 0.0135)
    return o, g


def test_ticks_match_oracle(oracle, gpu):
    o, g = seeded_pair(oracle, gpu)
    live0 = o.stats()["n_live"]
    for step in (1, 1, 3, 10, 25):
        o.step(step)
        g.step(step)
        compare_cells(o.download(), g.download(), exact=("id", "flags", "death_cause", "rng_state"))
        compare_tau(o.tau_download(), g.tau_download())
        assert o.tau_counters() == g.tau_counters()
    so, sg = o.stats(), g.stats()
    assert so["n_live"] < live0
    assert (so["n_taumoeba"], so["n_live"], so["n_dead"]) == (sg["n_taumoeba"], sg["n_live"], sg["n_dead"])
    assert so["mean_tau_tolerance"] == sg["mean_tau_tolerance"]
    t = g.tau_download()
    assert t["generation"].max() >= 1                           # divisions happened
    assert len(t["id"]) < 40 + (t["generation"] > 0).sum()      # and deaths were compacted away
    o.close()
    g.close()


def test_snapshot_carries_predators(gpu):
    w = crowded(gpu, ah.FP64_CONTRACT)
    w.step(20)
    data = w.snapshot()
    h = asph.parse(data, ah.CELL_FIELDS, ah.TAU_FIELDS)["header"]
    assert (h["taumoeba_count"], h["taumoeba_capacity"]) == (w.tau_counters()[0], 128)
    kw = dict(SMALL, co2_init=0.0, biology_rate=1e5, tau_compaction_enabled=1)
    with ah.World.from_snapshot(gpu, data, **kw) as b:
        assert b.tau_counters() == w.tau_counters()
        assert b.state_hash() == w.state_hash()
        w.step(15)
        b.step(15)
        assert b.state_hash() == w.state_hash()
    w.close()


def test_fp32_layout_reproducible(gpu):
    hashes, eaten = [], []
    for _ in range(2):
        with crowded(gpu, ah.FP32_HOT) as w:
            live0 = w.stats()["n_live"]
            w.step(60)
            hashes.append(w.state_hash())
            eaten.append(live0 - w.stats()["n_live"])
    assert hashes[0] == hashes[1] and eaten[0] > 0


def test_ensemble_refuses_predators(gpu):
    a = crowded(gpu, ah.FP32_HOT, cells=100, tau=2)
    b = crowded(gpu, ah.FP32_HOT, cells=100, tau=0)
    with pytest.raises(ah.AstroError) as e:
        ah.Ensemble(gpu, [a, b])
    assert e.value.status == ah.UNSUPPORTED
    a.close()
    b.close()
