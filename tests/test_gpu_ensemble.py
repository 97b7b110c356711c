"""Batched ensembles: chambers of uneven populations stepped together, each
stage one launch over every chamber and each step one graph launch.

The reference steps one world per call (src/sim/step.cu:109-184) and has no
ensemble, so the oracle here is the world path itself: every chamber of an
ensemble must end bit-identical (state hash over every SoA array and field,
plus the counters) to the same world stepped alone, which the world parity
tests pin against the CPU restatement.
"""
import numpy as np
import pytest

from conftest import ah, compare_cells

pytestmark = pytest.mark.gpu

# Small chambers with 64^3 temperature and irradiance grids, lit from a
# per-chamber in-plane direction, every cell awake (feed, taxis, light).
LIT = dict(chamber_w=550e-6, chamber_h=550e-6, chamber_d=100e-6, physics_rate=0.1, light_enabled=1,
           light_irradiance=1e4, grid={0: (64, 64, 64), 1: (16, 16, 16), 2: (-1, 0, 0), 3: (64, 64, 64)})
# 2D reference-shaped fields with CO2 rationing, division and compaction.
LIFE = dict(grid={0: (128, 128, 1), 1: (64, 64, 1), 2: (32, 32, 1), 3: (128, 128, 1)}, compaction_enabled=1,
            co2_init=2e-4, biology_rate=1e6)


def make_world(gpu, n, seed, kw, spawn):
    w = ah.World(gpu, capacity=2 * n + 64, **kw)
    w.spawn(n, seed, **spawn)
    return w


def assert_same(a: ah.World, b: ah.World, tag: str):
    ca, cb = a.counters(), b.counters()
    assert (ca.count, ca.next_id, ca.divisions_total, ca.deaths_total) == \
           (cb.count, cb.next_id, cb.divisions_total, cb.deaths_total), tag
    if a.state_hash() != b.state_hash():
        # name the first array that differs
        da, db = a.download(), b.download()
        compare_cells(da, db)
        for k in da:
            assert np.array_equal(da[k], db[k]), f"{tag}: {k}"
        for f in (ah.TEMPERATURE, ah.CO2, ah.N2):
            if a.field_dims(f)[0][0] > 0:
                assert np.array_equal(a.field(f), b.field(f)), f"{tag}: field {f}"
        raise AssertionError(f"{tag}: state hash differs")


def lit_chambers(gpu, counts, precision, seed0):
    rng = np.random.default_rng(seed0)
    out = []
    for i, n in enumerate(counts):
        ang = rng.uniform(0.0, 2.0 * np.pi)
        kw = dict(LIT, precision=precision, light_dir_x=float(np.cos(ang)), light_dir_y=float(np.sin(ang)),
                  seed=seed0 + i)
        spawn = dict(charge_dist=ah.DIST_UNIFORM, charge_a=0.0, charge_b=0.2, awake=True)
        out.append((n, seed0 + 100 + i, kw, spawn))
    return out


@pytest.mark.parametrize("precision", [ah.FP32_HOT, ah.FP64_CONTRACT])
def test_ensemble_matches_standalone_lit(gpu, precision):
    """Six lit chambers, 40 to 5,000 cells, each with its own light
    direction: graph-replayed and left-over eager ticks both match."""
    specs = lit_chambers(gpu, [40, 700, 2500, 150, 1800, 5000], precision, 11)
    ens_worlds = [make_world(gpu, *s) for s in specs]
    solo = [make_world(gpu, *s) for s in specs]
    with ah.Ensemble(gpu, ens_worlds) as e:
        assert len(e) == len(specs)
        for ticks in (5, 1, 2):
            e.step(ticks)
            for w in solo:
                w.step(ticks)
            e.sync()
            for i, (a, b) in enumerate(zip(ens_worlds, solo)):
                assert_same(a, b, f"chamber {i} after {ticks}")
    for w in ens_worlds + solo:
        w.close()


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_ensemble_matches_standalone_lifecycle(gpu, precision):
    """Crowded CO2-starved chambers that divide, die and compact every tick,
    with different populations and seeds."""
    specs = []
    for i, n in enumerate([300, 2200, 900, 60]):
        kw = dict(LIFE, precision=precision, seed=30 + i)
        spawn = dict(placement=ah.PLACE_GAUSSIAN, place_radius=60e-6, charge_dist=ah.DIST_UNIFORM, charge_b=0.9)
        specs.append((n, 40 + i, kw, spawn))
    ens_worlds, solo = [], []
    rng = np.random.default_rng(3)
    quota = 2.1e-14
    for s in specs:
        a, b = make_world(gpu, *s), make_world(gpu, *s)
        cells = a.download()
        n = len(cells["id"])
        cells["co2_held"] = rng.uniform(0.0, 1.0, n) * quota
        near = rng.random(n) < 0.1
        cells["co2_held"][near] = quota - rng.uniform(0.0, 4.0, near.sum()) * 6e-20
        dead = rng.random(n) < 0.05
        cells["flags"][dead] &= ~np.uint32(ah.FLAG_ALIVE)
        a.upload(cells)
        b.upload(cells)
        ens_worlds.append(a)
        solo.append(b)
    with ah.Ensemble(gpu, ens_worlds) as e:
        for _ in range(4):
            e.step(1)
            for w in solo:
                w.step(1)
        e.step(6)
        for w in solo:
            w.step(6)
        e.sync()
        for i, (a, b) in enumerate(zip(ens_worlds, solo)):
            assert_same(a, b, f"chamber {i}")
        assert sum(w.counters().divisions_total for w in solo) > 0
        assert sum(w.counters().deaths_total for w in solo) > 0
    for w in ens_worlds + solo:
        w.close()


def test_ensemble_profiled_steps(gpu):
    """Profiled (eager, per-stage timed) ensemble steps give the same state
    and time every batched stage."""
    specs = lit_chambers(gpu, [500, 3000, 90], ah.FP32_HOT, 70)
    ens_worlds = [make_world(gpu, *s) for s in specs]
    solo = [make_world(gpu, *s) for s in specs]
    with ah.Ensemble(gpu, ens_worlds) as e:
        e.profile(True)
        e.step(3)
        for w in solo:
            w.step(3)
        e.sync()
        prof = e.profile_read()
        for stage in ("hash_key", "thermal", "feed_walk", "contact_walk", "integrate_demand", "ftcs_T", "life_apply"):
            assert stage in prof and prof[stage][0] > 0, stage
        for i, (a, b) in enumerate(zip(ens_worlds, solo)):
            assert_same(a, b, f"chamber {i}")
    for w in ens_worlds + solo:
        w.close()


def test_ensemble_rejects_mismatched_chambers(gpu):
    a = ah.World(gpu, capacity=100, precision=ah.FP32_HOT, **LIT)
    b = ah.World(gpu, capacity=100, precision=ah.FP64_CONTRACT, **LIT)
    c = ah.World(gpu, capacity=100, precision=ah.FP32_HOT, **LIT)
    with pytest.raises(ah.AstroError) as err:
        ah.Ensemble(gpu, [a, b])
    assert err.value.status == ah.INVALID_ARGUMENT
    with pytest.raises(ah.AstroError):
        ah.Ensemble(gpu, [a, a])
    with ah.Ensemble(gpu, [a, c]) as e:
        c.set_clock(0.2, 1.0)   # a different physics clock: caught at the next step
        with pytest.raises(ah.AstroError) as err:
            e.step(1)
        assert err.value.status == ah.INVALID_ARGUMENT
    for w in (a, b, c):
        w.close()


def test_ensemble_autotune_keeps_results(gpu):
    """The autotuner's trial ticks are real ticks, and the launch shapes it
    picks change no result: chambers still match standalone worlds."""
    specs = lit_chambers(gpu, [1200, 80, 4000, 600], ah.FP32_HOT, 90)
    ens_worlds = [make_world(gpu, *s) for s in specs]
    solo = [make_world(gpu, *s) for s in specs]
    with ah.Ensemble(gpu, ens_worlds) as e:
        tuning = e.autotune(1)   # eleven trials of one tick
        assert "contact_walk" in tuning["stage_blocks"] and "ftcs_T" in tuning["ftcs_lanes"]
        for w in solo:
            w.step(11)
        e.sync()
        for i, (a, b) in enumerate(zip(ens_worlds, solo)):
            assert_same(a, b, f"chamber {i} after tuning")
        e.step(4)
        for w in solo:
            w.step(4)
        e.sync()
        for i, (a, b) in enumerate(zip(ens_worlds, solo)):
            assert_same(a, b, f"chamber {i} with tuned shapes")
    for w in ens_worlds + solo:
        w.close()


def test_ensemble_follows_chamber_changes(gpu):
    """A chamber changed between ensemble steps (its light, every chamber's
    clock) is stepped with the new values: the ensemble captures its step
    again instead of replaying the arguments it captured before."""
    specs = lit_chambers(gpu, [900, 2400, 300], ah.FP32_HOT, 130)
    ens_worlds = [make_world(gpu, *s) for s in specs]
    solo = [make_world(gpu, *s) for s in specs]
    with ah.Ensemble(gpu, ens_worlds) as e:
        e.step(4)
        for w in solo:
            w.step(4)
        e.sync()
        for i, (a, b) in enumerate(zip(ens_worlds, solo)):
            assert_same(a, b, f"chamber {i} before the change")
        h_before = [w.state_hash() for w in ens_worlds]
        for w in (ens_worlds[1], solo[1]):
            w.set_light(0.0, -1.0, 2.5e4, 5.0, True)
        for w in ens_worlds + solo:
            w.set_clock(0.2, 3.0)
        e.step(4)
        for w in solo:
            w.step(4)
        e.sync()
        for i, (a, b) in enumerate(zip(ens_worlds, solo)):
            assert_same(a, b, f"chamber {i} after the change")
            assert a.time()[1] == b.time()[1], f"chamber {i} clock"
        assert [w.state_hash() for w in ens_worlds] != h_before
    for w in ens_worlds + solo:
        w.close()


def test_ensemble_full_shape_seed7(gpu):
    """The batched-ensemble workload at its full shape and horizon: 256
    chambers of 1 x 1 x 0.2 mm, cell counts log-uniform in [500, 50,000] from
    seed 7 (about 2.76M cells in all), 64^3 temperature and irradiance grids
    each, a random in-plane light direction per chamber, every cell awake.
    1,000 ensemble steps (graph-replayed) against the same 256 chambers
    stepped one after another, each on its own graph: every chamber's state
    hash and counters are equal after 3 steps and after 1,000."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    import bench
    ens_worlds = bench.ensemble_chambers(gpu, 256)
    solo = bench.ensemble_chambers(gpu, 256)
    counts = [w.counters().count for w in solo]
    assert len(counts) == 256 and min(counts) >= 500 and max(counts) <= 50_000
    assert 2.0e6 < sum(counts) < 3.5e6
    with ah.Ensemble(gpu, ens_worlds) as e:
        for steps in (3, 997):
            e.step(steps)
            for w in solo:
                w.step(steps)
            e.sync()
            for i, (a, b) in enumerate(zip(ens_worlds, solo)):
                assert_same(a, b, f"chamber {i} (of {counts[i]} cells) after {steps}")
            print(f"256 chambers equal after {steps} more steps", flush=True)
    for w in ens_worlds + solo:
        w.close()
