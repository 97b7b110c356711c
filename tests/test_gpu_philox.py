"""The counter-based stream (rng_kind = ASTRO_RNG_PHILOX) on MI355X against
the oracle: the generator bit for bit (published known answers and random
counters), the per-cell streams, spawn, and whole ticks of the small plain
workload (2,000 cells, 64^3 temperature at 369.0 K, dt = 1e-4 s, seed 1)."""
import numpy as np
import pytest

from conftest import ah, compare_cells
from test_gpu_parity import FP32_POS_TOL, FP32_VEL_TOL, REF2D, rel_close
from test_philox_oracle import kat_words

pytestmark = pytest.mark.gpu


def test_device_philox_blocks(gpu, oracle):
    words, want = kat_words()
    assert np.array_equal(gpu.philox_blocks(words), want)
    rnd = np.random.default_rng(3).integers(0, 2 ** 32, (20_000, 6), dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(gpu.philox_blocks(rnd), oracle.philox_blocks(rnd))


def test_device_streams(gpu, oracle):
    rng = np.random.default_rng(8)
    for _ in range(40):
        seed = int(rng.integers(0, 2 ** 63))
        cid = int(rng.integers(1, 2 ** 40))
        tick = int(rng.integers(0, 2 ** 32))
        stage = int(rng.integers(1, 4))
        n = int(rng.integers(1, 40))
        assert np.array_equal(gpu.philox_stream(seed, cid, tick, stage, n),
                              oracle.philox_stream(seed, cid, tick, stage, n))


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_spawn_parity(gpu, oracle, precision):
    import bench
    kw = dict(rng_kind=ah.RNG_PHILOX)
    o = ah.World(oracle, bench.workload_c1(oracle, precision, **kw))
    g = ah.World(gpu, bench.workload_c1(gpu, precision, **kw))
    bench.populate_c1(o)
    bench.populate_c1(g)
    a, b = o.download(), g.download()
    compare_cells(a, b)
    for name in ("x", "y", "z"):
        want = a[name] if precision == ah.FP64_CONTRACT else a[name].astype(np.float32).astype(np.float64)
        assert np.array_equal(want, b[name]), name
    assert np.array_equal(a["energy"], b["energy"])
    assert not b["rng_state"].any()
    o.close()
    g.close()


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_small_workload_ticks(gpu, oracle, precision):
    """The small plain workload as specified (Philox seed 1), 20 ticks with
    thermal noise: ids, flags, death causes exact and rng_state 0; FP64
    positions within 1e-12 relative (libm ulps in the Gaussians); FP32 within
    FP32_POS_TOL with the oracle storing floats where the device does."""
    import bench
    fp32 = precision == ah.FP32_HOT
    o = ah.World(oracle, bench.workload_c1(oracle, precision, rng_kind=ah.RNG_PHILOX))
    g = ah.World(gpu, bench.workload_c1(gpu, precision, rng_kind=ah.RNG_PHILOX))
    bench.populate_c1(o)
    cells = o.download()
    if fp32:
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            cells[k] = cells[k].astype(np.float32).astype(np.float64)
        o.upload(cells)
        oracle.lib.oracle_world_set_float_store(o.h, 1)
    g.upload(cells)
    for ticks in (1, 19):
        o.step(ticks)
        g.step(ticks)
        a, b = o.download(), g.download()
        compare_cells(a, b)
        assert not b["rng_state"].any()
        if fp32:
            for name in ("x", "y", "z"):
                rel_close(a[name], b[name], 0.0, atol=FP32_POS_TOL, name=name)
            for name in ("vx", "vy", "vz"):
                rel_close(a[name], b[name], 0.0, atol=FP32_VEL_TOL, name=name)
        else:
            for name in ("x", "y", "z", "vx", "vy", "vz", "energy"):
                rel_close(a[name], b[name], 1e-12, name=name)
        rel_close(o.field(ah.TEMPERATURE), g.field(ah.TEMPERATURE), 1e-6, name="T")
    assert o.counters().tick == g.counters().tick == 20
    o.close()
    g.close()


def test_lit_tumbles(gpu, oracle):
    """Awake lit cells whose remembered signal exceeds what they see all
    tumble on the first tick (the taxis stream), then run: flags, death
    causes and the tumbled headings agree with the oracle."""
    kw = dict(REF2D, capacity=3000, chamber_w=1e-3, chamber_h=1e-3, ambient_temp=300.0, light_enabled=1,
              light_irradiance=1.0e4, co2_init=1e-3, biology_rate=1e6, rng_kind=ah.RNG_PHILOX, seed=12)
    o = ah.World(oracle, **kw)
    o.spawn(3000, 3, charge_dist=ah.DIST_UNIFORM, charge_b=0.9, awake=True)
    cells = o.download()
    cells["taxis_memory"][:] = 1.0e9
    o.upload(cells)
    g = ah.World(gpu, **kw)
    g.upload(cells)
    o.step(1)
    g.step(1)
    a, b = o.download(), g.download()
    compare_cells(a, b)
    assert np.count_nonzero(b["run_timer"] == 0.0) > 1000
    for name in ("dir_x", "dir_y", "dir_z"):
        rel_close(a[name], b[name], 1e-6, name=name)
    o.step(2)
    g.step(2)
    a, b = o.download(), g.download()
    compare_cells(a, b)
    for name in ("x", "y", "z", "energy"):
        rel_close(a[name], b[name], 1e-9, name=name)
    o.close()
    g.close()


def test_graph_eager_and_restore(gpu):
    """The device's tick counter drives the streams: graph-replayed ticks,
    eager (profiled) ticks and a world restored from a snapshot halfway all
    end in the same state."""
    kw = dict(REF2D, capacity=2500, chamber_w=1e-3, chamber_h=1e-3, light_enabled=1, light_irradiance=1e4,
              rng_kind=ah.RNG_PHILOX, precision=ah.FP32_HOT, seed=77)
    a, b = ah.World(gpu, **kw), ah.World(gpu, **kw)
    for w in (a, b):
        w.spawn(2000, 9, charge_dist=ah.DIST_UNIFORM, charge_b=0.9, awake=True)
    a.step(6)
    b.profile(True)
    for _ in range(6):
        b.step(1)
    b.profile(False)
    assert a.state_hash() == b.state_hash()
    with ah.World.from_snapshot(gpu, a.snapshot(), **kw) as r:
        assert r.counters().tick == 6
        a.step(5)
        r.step(5)
        assert a.state_hash() == r.state_hash()
    a.close()
    b.close()


def test_ensemble_of_counter_stream_chambers(gpu):
    from test_gpu_ensemble import assert_same, lit_chambers, make_world
    specs = [(n, s, dict(kw, rng_kind=ah.RNG_PHILOX), sp)
             for n, s, kw, sp in lit_chambers(gpu, [400, 3000, 1200], ah.FP32_HOT, 210)]
    ens = [make_world(gpu, *s) for s in specs]
    solo = [make_world(gpu, *s) for s in specs]
    mixed = make_world(gpu, *lit_chambers(gpu, [100], ah.FP32_HOT, 300)[0])
    with pytest.raises(ah.AstroError):
        ah.Ensemble(gpu, [ens[0], mixed])   # PCG32 and Philox chambers do not batch together
    with ah.Ensemble(gpu, ens) as e:
        e.step(4)
        for w in solo:
            w.step(4)
        e.sync()
        for i, (a, b) in enumerate(zip(ens, solo)):
            assert_same(a, b, f"chamber {i}")
    for w in ens + solo + [mixed]:
        w.close()
