"""The FP32 storage layout over long horizons on MI355X.

The per-tick parity tests hold the FP32 layout to the oracle over a few
ticks; these check that nothing drifts over seconds of simulated time, with
the physics the reference's own tests fix:
  * T4 Einstein diffusion (tests/physics/test_motion.cpp:90-113): 1 s of free
    dormant cells, MSD = 4 D t in the force-free plane, with either stream;
  * T1/T2 terminal velocities (test_motion.cpp:53-88): a full cell settles
    at 1.681 mm/s and an empty one rises at 52.1 um/s, noise off;
  * the reference's scenario acceptance gates (T24) run in the FP32 layout,
    each over its own horizon (up to 60 s and 60,000 ticks).
"""
import math

import pytest

from conftest import ah
from test_gpu_scenario import RUNNABLE, text_of
from test_oracle_golden import EX, K, T_ROOM, _drag

pytestmark = pytest.mark.gpu

FREE = dict(contact_enabled=0, thermal_enabled=0, emission_enabled=0, taxis_enabled=0, lifecycle_enabled=0,
            adhesion_enabled=0, precision=ah.FP32_HOT,
            grid={ah.TEMPERATURE: (16, 16, 1), ah.CO2: (16, 16, 1), ah.N2: (-1, 0, 0), ah.IRRADIANCE: (16, 16, 1)})


@pytest.mark.parametrize("rng", [ah.RNG_PCG32, ah.RNG_PHILOX])
def test_einstein_diffusion_over_one_second(gpu, oracle, rng):
    n = 4000
    with ah.World(gpu, capacity=n, rng_kind=rng, **FREE) as w:
        w.spawn(n, 20260802, charge_dist=ah.DIST_CONSTANT, charge_a=0.0)
        c0 = w.download()
        w.step(int(round(1.0 / w.dt)))
        c1 = w.download()
    t_local = float(__import__("numpy").float32(T_ROOM))   # the sampled (float) field value
    D = K["K_BOLTZ"] * t_local / _drag(oracle.lib, t_local)
    dx, dz = c1["x"] - c0["x"], c1["z"] - c0["z"]
    msd = float((dx * dx + dz * dz).mean())
    print(f"MSD {msd:.4e} m^2 vs 4 D t {4 * D:.4e} (ratio {msd / (4 * D):.4f})", flush=True)
    assert abs(msd / (4.0 * D) - 1.0) <= EX["T4_DIFFUSIVITY_20C_TOL"]
    assert abs(math.sqrt(msd) / EX["T4_RMS_DISP_1S_20C"] - 1.0) <= EX["T4_RMS_DISP_1S_20C_TOL"]


def test_terminal_velocities(gpu):
    """Noise off, 4 s: every full cell falls and every empty cell rises at
    the terminal velocity of the reference's expectations (0.5 %). The
    chamber is 4 cm tall so neither reaches a wall."""
    n = 1000
    with ah.World(gpu, capacity=2 * n, chamber_h=4e-2, thermal_noise=0, **FREE) as w:
        w.spawn(n, 3, placement=ah.PLACE_GAUSSIAN, place_radius=5e-4, charge_dist=ah.DIST_CONSTANT, charge_a=1.0)
        w.spawn(n, 4, placement=ah.PLACE_GAUSSIAN, place_radius=5e-4, charge_dist=ah.DIST_CONSTANT, charge_a=0.0)
        w.step(int(round(4.0 / w.dt)))
        c = w.download()
    vy = c["vy"]
    full, empty = vy[:n], vy[n:]
    assert full.max() < 0.0 and empty.min() > 0.0
    for v, want, tol in ((full, EX["T1_V_SETTLE_FULL"], EX["T1_V_SETTLE_FULL_TOL"]),
                         (empty, EX["T2_V_RISE_EMPTY"], EX["T2_V_RISE_EMPTY_TOL"])):
        assert abs(-v.mean() / want - 1.0) <= tol, (v.mean(), want)
        assert v.max() - v.min() <= 1e-4 * abs(want)   # every cell alike: no drift between them


@pytest.mark.parametrize("name", RUNNABLE)
def test_reference_acceptance_fp32(gpu, name):
    """Each reference scenario's accept block in the FP32 layout."""
    sc = ah.scenario_parse(gpu, text_of(name))
    with ah.scenario_world(gpu, sc, precision=ah.FP32_HOT) as w:
        assert w.desc.precision == ah.FP32_HOT
        res = ah.scenario_assert(gpu, sc, w)
        lines = [f"{c['metric']} {c['measured']:.6g} {c['op']} {c['value']:.6g} {'PASS' if c['passed'] else 'FAIL'}"
                 for c in res["checks"]]
        print(f"{name} (FP32): {res['ticks']} ticks, {res['sim_time_s']:.2f} s; " + "; ".join(lines), flush=True)
        assert res["all_pass"], lines
