"""The plain workloads over their whole specified horizons on MI355X, against
the oracle (its all-cores build, bit-identical to the serial one) stepped the
same number of steps from the same cells.

Small (2,000 cells in a 550 x 550 x 100 um slab, charge uniform in [0, 0.2],
a 64^3 temperature grid at 369.0 K, no light, dt = 1e-4 s, Philox seed 1),
1,000 steps: ids, flags, death causes, the counters and the temperature field
stay exact, and the kinematics stay within:
  * FP64 layout: 1e-12 relative (measured after 1,000 steps: 9e-19 m);
  * FP32 layout (the oracle storing floats where the device does): 1e-8 m in
    position and 1e-6 m/s in velocity. A coordinate's stored rounding can
    flip by one float ulp (3e-11 m at 275 um) on a step where the two sides'
    float and double displacement terms straddle a rounding boundary, and
    contacts carry such flips on; measured after 1,000 steps: 9.8e-10 m and
    5.1e-8 m/s, so the bounds hold ten times that, under 0.2 % of a cell
    radius.
"""
import numpy as np
import pytest

from conftest import ROOT, ah, compare_cells

pytestmark = pytest.mark.gpu

STEPS = 1000


@pytest.fixture(scope="module")
def omp_oracle():
    import subprocess
    lib = ROOT / "oracle" / "liboracle_omp.so"
    if not lib.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle"), "omp"], check=True, stdout=subprocess.DEVNULL)
    return ah.Binding(lib, "oracle_")


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_small_workload_full_horizon(gpu, omp_oracle, precision):
    import bench
    fp32 = precision == ah.FP32_HOT
    o = ah.World(omp_oracle, bench.workload_c1(omp_oracle, precision, rng_kind=ah.RNG_PHILOX))
    g = ah.World(gpu, bench.workload_c1(gpu, precision, rng_kind=ah.RNG_PHILOX))
    bench.populate_c1(o)
    cells = o.download()
    if fp32:
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            cells[k] = cells[k].astype(np.float32).astype(np.float64)
        o.upload(cells)
        omp_oracle.lib.oracle_world_set_float_store(o.h, 1)
    g.upload(cells)
    c = g.counters()
    c.next_id = o.counters().next_id   # an upload sets the count, not the next id
    g.set_counters(c)
    o.step(STEPS)
    g.step(STEPS)
    a, b = o.download(), g.download()
    compare_cells(a, b)
    ca, cb = o.counters(), g.counters()
    assert (ca.tick, ca.count, ca.next_id, ca.deaths_total) == (cb.tick, cb.count, cb.next_id, cb.deaths_total)
    dpos = max(float(np.abs(a[k] - b[k]).max()) for k in ("x", "y", "z"))
    dvel = max(float(np.abs(a[k] - b[k]).max()) for k in ("vx", "vy", "vz"))
    dT = float(np.abs(o.field(ah.TEMPERATURE).astype(np.float64) - g.field(ah.TEMPERATURE)).max())
    print(f"after {STEPS} steps: max |dx| {dpos:.3e} m, max |dv| {dvel:.3e} m/s, max |dT| {dT:.3e} K", flush=True)
    assert dT == 0.0
    if fp32:
        assert dpos <= 1e-8 and dvel <= 1e-6, (dpos, dvel)
    else:
        for k in ("x", "y", "z", "vx", "vy", "vz", "energy"):
            scale = float(np.abs(a[k]).max()) or 1.0
            assert float(np.abs(a[k] - b[k]).max()) <= 1e-12 * scale, k
    o.close()
    g.close()


def _full_run(gpu, oracle, make, fill):
    o = ah.World(oracle, make(oracle))
    g = ah.World(gpu, make(gpu))
    fill(o)
    cells = o.download()
    g.upload(cells)
    c = g.counters()
    c.next_id = o.counters().next_id
    g.set_counters(c)
    return o, g


def _max_diff(a, b, names):
    return max(float(np.abs(a[k] - b[k]).max()) for k in names)


def test_real_shape_exact_then_statistical(gpu, omp_oracle):
    """Real shape (25,000 awake cells, 256 x 256 x 64 T + nutrient with a hot
    spot, light along +x), FP64 layout, 2,000 steps. The lit, swimming,
    wall-sticking population is chaotic: the oracle's glibc and the device's
    OCML differ by ulps in exp, log, sin and cos, and the difference
    grows about 70-fold per 100 steps (measured max |dx|: 7e-18 m at step
    100, 2e-15 at 200, 6e-10 at 500, then whole cells by 1,000). So:
      * at step 300 every integer (ids, flags, death causes, streams) is
        exact and positions agree within 1e-9 m;
      * at step 2,000 the two populations agree in distribution: the same
        count, the fraction of cells stuck to a wall within 1 % of the
        population, and two-sample Kolmogorov-Smirnov tests of height and
        speed that do not reject at the 1e-3 level."""
    import bench
    from scipy.stats import ks_2samp
    o, g = _full_run(gpu, omp_oracle, lambda b: bench.workload_c2(b, ah.FP64_CONTRACT), bench.populate_c2)
    o.step(300)
    g.step(300)
    a, b = o.download(), g.download()
    compare_cells(a, b)
    d300 = _max_diff(a, b, ("x", "y", "z"))
    print(f"C2 step 300: max |dx| {d300:.3e} m", flush=True)
    assert d300 <= 1e-9
    o.step(1700)
    g.step(1700)
    a, b = o.download(), g.download()
    n = len(a["flags"])
    assert n == len(b["flags"]) and o.counters().count == g.counters().count
    stuck = [int(((c["flags"] & ah.FLAG_STUCK) != 0).sum()) for c in (a, b)]
    speed = [np.sqrt(c["vx"] ** 2 + c["vy"] ** 2 + c["vz"] ** 2) for c in (a, b)]
    p_z = ks_2samp(a["z"], b["z"]).pvalue
    p_v = ks_2samp(speed[0], speed[1]).pvalue
    print(f"C2 step 2000: stuck {stuck}, KS p height {p_z:.3g}, speed {p_v:.3g}", flush=True)
    assert abs(stuck[0] - stuck[1]) <= 0.01 * n
    assert p_z > 1e-3 and p_v > 1e-3
    o.close()
    g.close()


def test_scale_full_horizon_exact(gpu, omp_oracle):
    """The 1M-cell workload (division and compaction every step), FP64
    layout, 500 steps: dormant cells in an unlit chamber keep the chaos of the
    lit case out, and every integer stays exact to the end -- ids, flags,
    death causes, streams, the population, next id, divisions and deaths --
    with positions within 1e-12 relative (measured: 7e-18 m)."""
    import bench
    o, g = _full_run(gpu, omp_oracle, lambda b: bench.workload(b, 1_000_000, ah.FP64_CONTRACT, ah.RNG_PHILOX),
                     lambda w: bench.populate(w, 1_000_000))
    o.step(500)
    g.step(500)
    a, b = o.download(), g.download()
    ca, cb = o.counters(), g.counters()
    print(f"C3 after 500 steps: count {ca.count}/{cb.count}, divisions {ca.divisions_total}/{cb.divisions_total}, "
          f"deaths {ca.deaths_total}/{cb.deaths_total}", flush=True)
    assert (ca.count, ca.next_id, ca.divisions_total, ca.deaths_total) == \
           (cb.count, cb.next_id, cb.divisions_total, cb.deaths_total)
    assert ca.divisions_total > 0 and ca.deaths_total > 0
    compare_cells(a, b)
    for k in ("x", "y", "z", "vx", "vy", "vz", "energy"):
        scale = float(np.abs(a[k]).max()) or 1.0
        assert float(np.abs(a[k] - b[k]).max()) <= 1e-12 * scale, k
    o.close()
    g.close()
