"""The all-cores build of the oracle (oracle/liboracle_omp.so, OpenMP) is
the single-threaded oracle, bit for bit: its parallel loops write only their
own cell or voxel, or add integers into fixed-point accumulators. Checked on
the state hash and counters of worlds that exercise every parallel loop:
the thermostat, light and shadow walk, taxis, motion, rationed uptake,
division and compaction, on 2D and 3D grids, with both streams."""
import subprocess

import numpy as np
import pytest

from conftest import ROOT, ah

OMP = ROOT / "oracle" / "liboracle_omp.so"


@pytest.fixture(scope="module")
def omp():
    if not OMP.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle"), "omp"], check=True, stdout=subprocess.DEVNULL)
    return ah.Binding(OMP, "oracle_")


@pytest.mark.parametrize("case", ["scale", "lit3d", "philox2d"])
def test_omp_build_equals_serial(oracle, omp, case):
    import bench
    if case == "scale":
        make = lambda b: ah.World(b, bench.workload(b, 60_000, ah.FP64_CONTRACT))  # noqa: E731
        fill = lambda w: bench.populate(w, 60_000)  # noqa: E731
        ticks = 4
    elif case == "lit3d":
        make = lambda b: ah.World(b, bench.workload_c2(b, ah.FP32_HOT, cells=6000))  # noqa: E731
        fill = lambda w: bench.populate_c2(w, 6000)  # noqa: E731
        ticks = 3
    else:
        kw = dict(capacity=6000, chamber_w=1e-3, chamber_h=1e-3, light_enabled=1, light_irradiance=1e4,
                  rng_kind=ah.RNG_PHILOX, co2_init=2e-4, biology_rate=1e6, compaction_enabled=1,
                  boundary_x=ah.BOUNDARY_ABSORBING,
                  grid={0: (128, 128, 1), 1: (64, 64, 1), 2: (32, 32, 1), 3: (128, 128, 1)})
        make = lambda b: ah.World(b, **kw)  # noqa: E731
        fill = lambda w: w.spawn(4000, 3, placement=ah.PLACE_GAUSSIAN, place_radius=1.5e-4,
                                 charge_dist=ah.DIST_UNIFORM, charge_b=0.9, awake=True)  # noqa: E731
        ticks = 6
    a, b = make(oracle), make(omp)
    fill(a)
    fill(b)
    a.step(ticks)
    b.step(ticks)
    ca, cb = a.counters(), b.counters()
    assert (ca.count, ca.next_id, ca.divisions_total, ca.deaths_total) == \
           (cb.count, cb.next_id, cb.divisions_total, cb.deaths_total)
    if a.state_hash() != b.state_hash():
        da, db = a.download(), b.download()
        for k in da:
            assert np.array_equal(da[k], db[k]), k
        raise AssertionError("fields differ")
    if case == "scale":
        assert ca.divisions_total > 0 and ca.deaths_total > 0
    a.close()
    b.close()
