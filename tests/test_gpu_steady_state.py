"""T29: no device allocation in the steady-state tick loop
(tests/physics/test_perf.cu:28-74).

A world that divides and compacts every tick drives the lifecycle scans and
the compaction, the paths most likely to allocate. After 25 warm-up ticks
(which capture the tick's graph), free device memory may not fall by 8 MB or
more over 500 further ticks, while the population changes.
"""
import ctypes as C

import pytest

from conftest import ah

pytestmark = pytest.mark.gpu


def free_device_memory() -> int:
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_no_allocation_in_steady_state(gpu, precision):
    # The reference's set-up: capacity 2^20, saturated CO2 (cells divide),
    # absorbing side walls (cells die there) with compaction on, biology
    # clock 2e7, 40,000 dormant cells of charge 0.5 placed uniformly.
    w = ah.World(gpu, capacity=1 << 20, seed=20260802, co2_init=1.5, precision=precision,
                 boundary_x=ah.BOUNDARY_ABSORBING, boundary_y=ah.BOUNDARY_ABSORBING, compaction_enabled=1,
                 biology_rate=2.0e7)
    w.spawn(40_000, 20260802, charge_dist=ah.DIST_CONSTANT, charge_a=0.5)
    for _ in range(25):
        w.step(1)
    w.sync()
    before = free_device_memory()
    pop_before = w.counters().count
    for _ in range(500):
        w.step(1)
    w.sync()
    after = free_device_memory()
    c = w.counters()
    dropped = before - after
    print(f"T29: free {before / 2**20:.1f} -> {after / 2**20:.1f} MB over 500 ticks; "
          f"pop {pop_before} -> {c.count}, divisions {c.divisions_total}, deaths {c.deaths_total}", flush=True)
    assert c.count != pop_before
    assert c.divisions_total > 0 and c.deaths_total > 0
    assert dropped < (8 << 20)
    w.close()
