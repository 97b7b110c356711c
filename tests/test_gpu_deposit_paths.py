"""The thermostat's two deposit paths give the same world bit for bit.

thermal_kernel (src/sim/thermal.cu:19-71) adds every cell's exchange into the
temperature field's fixed-point accumulator and flush_kernel
(src/fields/diffuse.cu:51-60) folds it into the value. The device does that
on the ATOMIC path; on the DIRECT path a cell that is the only entry in its
voxel adds its own exchange to the value, which is the flush of a single
depositor. The world parity tests pin the default (automatic) choice against
the oracle; these pin the two forced paths against each other on chambers
where both alone and shared voxels, dormant cells and starving cells occur.
"""
import numpy as np
import pytest

from conftest import ah
from test_gpu_ensemble import assert_same

pytestmark = pytest.mark.gpu

GRID = {ah.TEMPERATURE: (64, 64, 64), ah.CO2: (16, 16, 4), ah.N2: (-1, 0, 0), ah.IRRADIANCE: (64, 64, 64)}


def chamber(gpu, precision, path, seed=5, n=12000, grid=GRID):
    ang = 0.7 + seed
    w = ah.World(gpu, capacity=2 * n + 4096, seed=seed, chamber_w=1e-3, chamber_h=1e-3, chamber_d=2e-4,
                 precision=precision, grid=grid, physics_rate=0.1, light_enabled=1, light_irradiance=1e4,
                 light_dir_x=float(np.cos(ang)), light_dir_y=float(np.sin(ang)))
    # spread awake cells (mostly alone in their voxel), a dense awake cluster
    # (shared voxels), dormant cells, and awake cells with so little energy
    # left that they starve within the first ticks' thermostat substeps (an
    # awake cell gives up about 5e-6 J per tick here)
    w.spawn(n, seed + 1, charge_dist=ah.DIST_LOG_UNIFORM, charge_a=1e-3, charge_b=0.2, awake=True)
    w.spawn(n // 4, seed + 2, placement=ah.PLACE_GAUSSIAN, place_radius=2e-5, charge_dist=ah.DIST_UNIFORM,
            charge_a=0.0, charge_b=0.2, awake=True)
    w.spawn(n // 8, seed + 3, charge_dist=ah.DIST_CONSTANT, charge_a=0.1)
    w.spawn(n // 16, seed + 4, charge_dist=ah.DIST_CONSTANT, charge_a=0.1, awake=True)
    cells = w.download()
    k = n // 16
    cells["energy"][-k:] = np.linspace(1e-9, 2e-5, k)
    w.upload(cells)
    if path is not None:
        w.set_deposit_path(path)
    return w


@pytest.mark.parametrize("precision", [ah.FP32_HOT, ah.FP64_CONTRACT])
def test_forced_paths_agree(gpu, precision):
    worlds = {p: chamber(gpu, precision, p) for p in (ah.DEPOSIT_ATOMIC, ah.DEPOSIT_DIRECT, ah.DEPOSIT_AUTO)}
    try:
        for ticks in (1, 6, 3):
            for w in worlds.values():
                w.step(ticks)
            a = worlds[ah.DEPOSIT_ATOMIC]
            for p in (ah.DEPOSIT_DIRECT, ah.DEPOSIT_AUTO):
                assert_same(a, worlds[p], f"path {p} after {ticks}")
        s, c = a.stats(), a.counters()
        print(f"live {s['n_live']} awake {s['n_awake']} deaths {c.deaths_total}", flush=True)
        assert c.deaths_total > 0 and s["n_awake"] > 0   # starvation and exchange both happened
    finally:
        for w in worlds.values():
            w.close()


def test_path_switch_between_ticks(gpu):
    """Switching the path between ticks (the captured tick is dropped and
    recaptured) leaves nothing stale: the counts are cleared each tick."""
    a = chamber(gpu, ah.FP32_HOT, ah.DEPOSIT_ATOMIC, seed=9)
    b = chamber(gpu, ah.FP32_HOT, ah.DEPOSIT_DIRECT, seed=9)
    try:
        for path, ticks in ((ah.DEPOSIT_ATOMIC, 2), (ah.DEPOSIT_DIRECT, 3), (ah.DEPOSIT_AUTO, 2),
                            (ah.DEPOSIT_DIRECT, 1)):
            b.set_deposit_path(path)
            a.step(ticks)
            b.step(ticks)
            assert_same(a, b, f"after switching to {path}")
    finally:
        a.close()
        b.close()


def test_ensemble_direct_matches_atomic_worlds(gpu):
    """An ensemble follows its first chamber's path: forced DIRECT there, the
    chambers match standalone worlds forced ATOMIC."""
    ens = [chamber(gpu, ah.FP32_HOT, ah.DEPOSIT_DIRECT if i == 0 else None, seed=20 + i, n=n)
           for i, n in enumerate((3000, 9000, 600))]
    solo = [chamber(gpu, ah.FP32_HOT, ah.DEPOSIT_ATOMIC, seed=20 + i, n=n) for i, n in enumerate((3000, 9000, 600))]
    try:
        with ah.Ensemble(gpu, ens) as e:
            for ticks in (4, 1):
                e.step(ticks)
                for w in solo:
                    w.step(ticks)
                e.sync()
                for i, (x, y) in enumerate(zip(ens, solo)):
                    assert_same(x, y, f"chamber {i} after {ticks}")
    finally:
        for w in ens + solo:
            w.close()


def test_unknown_path_rejected(gpu):
    w = chamber(gpu, ah.FP32_HOT, None, n=100)
    try:
        with pytest.raises(Exception, match="deposit path"):
            w.set_deposit_path(3)
    finally:
        w.close()
