"""Pins the CPU oracle to the reference's published answers, without the reference.

Two fixtures under tests/golden (written by tests/golden/make_golden.py):
  * ref_vectors.json: input/output vectors of the reference's own host-callable
    physics, compiled from its headers. The oracle must reproduce every output
    bit for bit.
  * expected_values.json: the reference's first-principles physics expectations
    T1-T26 (tests/golden/expected_values.h). The oracle must reproduce each one
    the way the reference's own tests check it:
    - tests/physics/test_motion.cpp:53-192 for T1-T4, T6, T8, the OU shape
      function and momentum relaxation;
    - test_thermal.cu:84-86 for T7;
    - test_taxis.cu:72-157 for T26;
    - test_lifecycle.cu:94 for T18;
    - test_fields.cu:76-156 for FTCS conservation and Gaussian spreading.
The closed forms are computed here from the canon constants, independently of
the oracle. A dropped term, a wrong sign or a transposed argument in the oracle
fails at least one of these.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT, ah

GOLD = Path(__file__).resolve().parent / "golden"
U64, U32, D, F, I = C.c_uint64, C.c_uint32, C.c_double, C.c_float, C.c_int32
PU64, PD = C.POINTER(C.c_uint64), C.POINTER(C.c_double)


def canon() -> dict:
    out = {}
    pat = re.compile(r"#define\s+ASTRO_(\w+)\s+([-+0-9.eE]+)\s*$")
    for line in (ROOT / "include" / "astro_canon.h").read_text().splitlines():
        m = pat.match(line.strip())
        if m:
            out[m.group(1)] = float(m.group(2))
    return out


K = canon()
EX = json.loads((GOLD / "expected_values.json").read_text())["values"]
VEC = json.loads((GOLD / "ref_vectors.json").read_text())["vectors"]
T_ROOM = K["AMBIENT_TEMP_DEFAULT"]


def fx(s: str) -> float:
    return float.fromhex(s)


def sig(lib, name, res, args):
    f = getattr(lib, "oracle_" + name)
    f.restype = res
    f.argtypes = args
    return f


@pytest.fixture(scope="module")
def lib(oracle):
    return oracle.lib


def close(got, want, rel):
    return abs(got - want) <= rel * abs(want)


# ---------------------------------------------------------------------------
# Reference vectors, bit for bit
# ---------------------------------------------------------------------------
def test_integer_streams_match_reference_vectors(lib):
    sm = sig(lib, "splitmix64", U64, [U64])
    for x, y in VEC["splitmix64"]:
        assert sm(int(x)) == int(y)
    ci = sig(lib, "cell_stream_inc", U64, [U64])
    for x, y in VEC["cell_stream_inc"]:
        assert ci(int(x)) == int(y)
    cr = sig(lib, "cell_rng_init", U64, [U64, U64])
    for s, i, y in VEC["cell_rng_init"]:
        assert cr(int(s), int(i)) == int(y)
    ps = sig(lib, "pcg_split", U64, [U64, U64])
    for s, i, y in VEC["pcg_split"]:
        assert ps(int(s), int(i)) == int(y)
    tf = sig(lib, "to_fixed", C.c_int64, [D, D])
    for x, sc, y in VEC["to_fixed"]:
        assert tf(fx(x), fx(sc)) == int(y)


def test_random_draws_match_reference_vectors(lib):
    ud = sig(lib, "uniform01d", D, [PU64, U64])
    gp = sig(lib, "gaussian_pair", None, [PU64, U64, PD, PD])
    for s in VEC["streams"]:
        st = U64(int(s["state"]))
        inc = int(s["inc"])
        assert [ud(C.byref(st), inc) for _ in s["uniform01d"]] == [fx(v) for v in s["uniform01d"]]
        for z0, z1 in s["gaussian_pairs"]:
            a, b = D(), D()
            gp(C.byref(st), inc, C.byref(a), C.byref(b))
            assert (a.value, b.value) == (fx(z0), fx(z1))
        assert st.value == int(s["final_state"])
    tt = sig(lib, "taxis_tumble_angle", D, [PU64, U64])
    for s in VEC["taxis_tumble_angle"]:
        st = U64(int(s["state"]))
        assert [tt(C.byref(st), int(s["inc"])) for _ in s["angles"]] == [fx(v) for v in s["angles"]]
        assert st.value == int(s["final_state"])
    da = sig(lib, "division_axis", None, [U64, PD, PD, PD])
    for did, ax, ay, az in VEC["division_axis"]:
        a, b, c = D(), D(), D()
        da(int(did), C.byref(a), C.byref(b), C.byref(c))
        assert (a.value, b.value, c.value) == (fx(ax), fx(ay), fx(az))


def test_motion_physics_matches_reference_vectors(lib):
    sh = sig(lib, "ou_position_shape", D, [D])
    for x, y in VEC["ou_position_shape"]:
        assert sh(fx(x)) == fx(y)
    wv = sig(lib, "water_viscosity", D, [D])
    for t, y in VEC["water_viscosity"]:
        assert wv(fx(t)) == fx(y)
    oc = sig(lib, "ou_coefficients", None, [D, D, D, D, D * 6])
    oa = sig(lib, "ou_advance_axis", None, [PD, PD, D, D, D * 6, D, I, D, D])
    for case in VEC["ou"]:
        mass, gamma, t, dt = map(fx, case["in"])
        c6 = (D * 6)()
        oc(mass, gamma, t, dt, c6)
        assert list(c6) == [fx(v) for v in case["coeff"]]
        p0, v0, f, z1, z2 = map(fx, case["advance_in"][:5])
        p, v = D(p0), D(v0)
        oa(C.byref(p), C.byref(v), f, gamma, c6, dt, case["advance_in"][5], z1, z2)
        assert (p.value, v.value) == tuple(map(fx, case["advance_out"]))


def test_light_heat_taxis_match_reference_vectors(lib):
    do = sig(lib, "disc_overlap_fraction", D, [D, D])
    for p, a, y in VEC["disc_overlap_fraction"]:
        assert do(fx(p), fx(a)) == fx(y)
    sf = sig(lib, "shadow_fraction", D, [D * 3, D * 3, D * 3])
    for s3, o3, l3, y in VEC["shadow_fraction"]:
        args = [(D * 3)(*map(fx, v)) for v in (s3, o3, l3)]
        assert sf(*args) == fx(y)
    le = sig(lib, "lumped_exchange_energy", D, [D, D, D, D, D])
    for *args, y in VEC["lumped_exchange_energy"]:
        assert le(*map(fx, args)) == fx(y)
    te = sig(lib, "taxis_ema_update", F, [F, D, D])
    for e0, s, dt, y in VEC["taxis_ema_update"]:
        assert te(fx(e0), fx(s), fx(dt)) == fx(y)


def test_predator_helpers_match_reference_vectors(lib):
    """tau_drag, tau_overlaps_cell, tau_should_tumble, tau_n2_hazard,
    tau_n2_death_prob, tau_daughter_tolerance (src/sim/predation.cuh:24-80),
    bit for bit, including overlaps within an ulp of contact and clamped
    tolerances."""
    td = sig(lib, "tau_drag", D, [D])
    for t, y in VEC["tau_drag"]:
        assert td(fx(t)) == fx(y)
    to = sig(lib, "tau_overlaps_cell", C.c_int32, [D * 3, D * 3])
    for a3, b3, y in VEC["tau_overlaps_cell"]:
        assert to((D * 3)(*map(fx, a3)), (D * 3)(*map(fx, b3))) == y
    tt = sig(lib, "tau_should_tumble", C.c_int32, [F, C.POINTER(F), F, D])
    for s_, e0, run, dt, y, e1 in VEC["tau_should_tumble"]:
        e = F(fx(e0))
        assert tt(fx(s_), C.byref(e), fx(run), fx(dt)) == y and e.value == fx(e1)
    th = sig(lib, "tau_n2_hazard", D, [D, D])
    for n2, tol, y in VEC["tau_n2_hazard"]:
        assert th(fx(n2), fx(tol)) == fx(y)
    tp = sig(lib, "tau_n2_death_prob", D, [D, D])
    for h, dtb, y in VEC["tau_n2_death_prob"]:
        assert tp(fx(h), fx(dtb)) == fx(y)
    tdt = sig(lib, "tau_daughter_tolerance", F, [F, D])
    for par, m, y in VEC["tau_daughter_tolerance"]:
        assert tdt(fx(par), fx(m)) == fx(y)


def test_hash_keys_match_reference_vectors(oracle):
    # One cell per world; its bucket is the only non-empty one.
    for nx, ny, nz, cell, x, y, z, key in VEC["hash_key"][:12]:
        cell = fx(cell)
        desc = dict(capacity=4, chamber_w=nx * cell * (1 - 1e-12), chamber_h=ny * cell * (1 - 1e-12),
                    chamber_d=nz * cell * (1 - 1e-12), hash_cell_size=cell,
                    grid={ah.TEMPERATURE: (8, 8, 1), ah.CO2: (8, 8, 1), ah.N2: (-1, 0, 0),
                          ah.IRRADIANCE: (8, 8, 1)})
        with ah.World(oracle, **desc) as w:
            assert w.hash_dims()[:3] == (nx, ny, nz)
            cells = {"id": np.array([1], np.uint64), "flags": np.array([ah.FLAG_OCCUPIED | ah.FLAG_ALIVE], np.uint32),
                     "x": np.array([fx(x)]), "y": np.array([fx(y)]), "z": np.array([fx(z)])}
            w.upload(cells)
            w.hash_build()
            _, bs, be = w.hash_download()
            assert list(np.nonzero(be > bs)[0]) == [key]


# ---------------------------------------------------------------------------
# The reference's physics expectations (T1-T26)
# ---------------------------------------------------------------------------
def _drag(lib, t):
    return 6.0 * math.pi * sig(lib, "water_viscosity", D, [D])(t) * K["CELL_RADIUS"]


def _mass(charge):
    return K["CELL_MASS_DRY"] + charge * K["CELL_ENERGY_MAX"] / (K["C_LIGHT"] ** 2)


def _buoyant_weight(m):
    return -(m - K["WATER_DENSITY"] * K["CELL_VOLUME"]) * K["G_ACCEL"]


def _terminal(lib, mass, force, seconds=4.0):
    """tests/physics/test_motion.cpp:28-41: noise off, one axis."""
    gamma = _drag(lib, T_ROOM)
    oc = sig(lib, "ou_coefficients", None, [D, D, D, D, D * 6])
    oa = sig(lib, "ou_advance_axis", None, [PD, PD, D, D, D * 6, D, I, D, D])
    c6 = (D * 6)()
    oc(mass, gamma, T_ROOM, K["DT_PHYSICS"], c6)
    p, v = D(0.0), D(0.0)
    for _ in range(int(seconds / K["DT_PHYSICS"])):
        oa(C.byref(p), C.byref(v), force, gamma, c6, K["DT_PHYSICS"], 0, 0.0, 0.0)
    return v.value


def test_t1_t2_t3_terminal_velocities(lib):
    vy_full = _terminal(lib, _mass(1.0), _buoyant_weight(_mass(1.0)))
    vy_empty = _terminal(lib, _mass(0.0), _buoyant_weight(_mass(0.0)))
    assert vy_full < 0.0 and close(-vy_full, EX["T1_V_SETTLE_FULL"], EX["T1_V_SETTLE_FULL_TOL"])
    assert vy_empty > 0.0 and close(-vy_empty, EX["T2_V_RISE_EMPTY"], EX["T2_V_RISE_EMPTY_TOL"])
    # neutral buoyancy: m_dry + q E_max / c^2 = rho V
    q = (K["WATER_DENSITY"] * K["CELL_VOLUME"] - K["CELL_MASS_DRY"]) * K["C_LIGHT"] ** 2 / K["CELL_ENERGY_MAX"]
    assert close(q, EX["T3_CHARGE_NEUTRAL"], EX["T3_CHARGE_NEUTRAL_TOL"])
    assert abs(_terminal(lib, _mass(q), _buoyant_weight(_mass(q)))) < 1e-3 * EX["T1_V_SETTLE_FULL"]
    assert _terminal(lib, _mass(0.5 * q), _buoyant_weight(_mass(0.5 * q))) > 0.0
    assert _terminal(lib, _mass(2.0 * q), _buoyant_weight(_mass(2.0 * q))) < 0.0


def test_t4_einstein_diffusion_through_the_step(oracle, lib):
    gamma = _drag(lib, T_ROOM)
    D_room = K["K_BOLTZ"] * T_ROOM / gamma
    assert close(D_room, EX["T4_DIFFUSIVITY_20C"], 1e-9)
    # 4000 empty cells, 1 s of the whole oracle step with noise on; x and z
    # carry no force (gravity is along y), so their 2D MSD is 4 D t.
    n = 4000
    with ah.World(oracle, capacity=n, contact_enabled=0, thermal_enabled=0, emission_enabled=0, taxis_enabled=0,
                  lifecycle_enabled=0, adhesion_enabled=0,
                  grid={ah.TEMPERATURE: (16, 16, 1), ah.CO2: (16, 16, 1), ah.N2: (-1, 0, 0),
                        ah.IRRADIANCE: (16, 16, 1)}) as w:
        w.spawn(n, 20260802, charge_dist=ah.DIST_CONSTANT, charge_a=0.0)
        c0 = w.download()
        w.step(int(round(1.0 / w.dt)))
        c1 = w.download()
    t_local = float(np.float32(T_ROOM))   # dormant cells drag at the sampled (float) field value
    D_run = K["K_BOLTZ"] * t_local / _drag(lib, t_local)
    msd = float(np.mean((c1["x"] - c0["x"]) ** 2 + (c1["z"] - c0["z"]) ** 2))
    assert close(msd, 4.0 * D_run * 1.0, EX["T4_DIFFUSIVITY_20C_TOL"])
    assert close(math.sqrt(msd), EX["T4_RMS_DISP_1S_20C"], EX["T4_RMS_DISP_1S_20C_TOL"])


def test_ou_shape_function_and_sub_einstein_variance(lib):
    sh = sig(lib, "ou_position_shape", D, [D])
    for x in (1e-8, 1e-6, 1e-4, 1e-3, 5e-3):
        assert close(sh(x), (2.0 / 3.0) * x ** 3, 1e-2) and sh(x) > 0.0
    assert close(sh(1e-2 * (1 - 1e-12)), sh(1e-2 * (1 + 1e-12)), 1e-6)
    assert close(sh(50.0), 2.0 * 50.0 - 3.0, 1e-12)
    gamma = _drag(lib, T_ROOM)
    m = _mass(1.0)
    beta = gamma / m
    x = beta * K["DT_PHYSICS"]
    c6 = (D * 6)()
    sig(lib, "ou_coefficients", None, [D, D, D, D, D * 6])(m, gamma, T_ROOM, K["DT_PHYSICS"], c6)
    var_direct = (K["K_BOLTZ"] * T_ROOM / m) / beta ** 2 * (2 * x - 3 + 4 * math.exp(-x) - math.exp(-2 * x))
    sigma_rr = c6[3]   # (e1, 1 - e1, 1/beta, sigma_rr, sigma_vv, sigma_rv)
    assert close(sigma_rr ** 2, var_direct, 1e-9)
    assert sigma_rr ** 2 < 2.0 * K["K_BOLTZ"] * T_ROOM / gamma * K["DT_PHYSICS"]


def test_t5_t6_t8_thrust_mass_reynolds(lib):
    assert close(K["CELL_ENERGY_MAX"] / K["C_LIGHT"] ** 2, EX["T5_DELTA_MASS_FULL"], EX["T5_DELTA_MASS_FULL_TOL"])
    thrust = 1.0e-3 / K["C_LIGHT"]
    assert close(thrust, EX["T6_THRUST_1MW"], EX["T6_THRUST_1MW_TOL"])
    # driven end to end in zero gravity: recoil along -x
    vx = _terminal(lib, _mass(0.0), -thrust)
    assert vx < 0.0 and close(-vx, EX["T6_VTERM_1MW"], EX["T6_VTERM_1MW_TOL"])
    eta = sig(lib, "water_viscosity", D, [D])(T_ROOM)
    re = K["WATER_DENSITY"] * EX["T1_V_SETTLE_FULL"] * K["CELL_DIAMETER"] / eta
    assert close(re, EX["T8_REYNOLDS_FULL"], EX["T8_REYNOLDS_FULL_TOL"])
    hover = -_buoyant_weight(_mass(1.0)) * K["C_LIGHT"]
    assert close(hover, EX["HOVER_POWER_FULL"], EX["HOVER_POWER_FULL_TOL"])
    assert hover < K["PETROVA_MAX_POWER"]


def test_momentum_relaxation_and_motility(lib):
    gamma = _drag(lib, T_ROOM)
    assert close(_mass(0.0) / gamma, EX["TAU_MOMENTUM_EMPTY"], EX["TAU_MOMENTUM_EMPTY_TOL"])
    assert close(_mass(1.0) / gamma, EX["TAU_MOMENTUM_FULL"], EX["TAU_MOMENTUM_FULL_TOL"])
    t_set = K["CELL_TEMP_SETPOINT"]
    d_set = K["K_BOLTZ"] * t_set / _drag(lib, t_set)
    d_room = K["K_BOLTZ"] * T_ROOM / gamma
    assert close(d_set, EX["T12_DIFFUSIVITY_SETPOINT"], EX["T12_DIFFUSIVITY_SETPOINT_TOL"])
    assert close(d_set / d_room, EX["T12_MOTILITY_RATIO"], EX["T12_MOTILITY_RATIO_TOL"])


def test_t7_conduction_through_lumped_exchange(lib):
    le = sig(lib, "lumped_exchange_energy", D, [D, D, D, D, D])
    dt = 1e-3
    # a voxel far larger than the cell's exchange leaves the linear law
    p = le(K["CELL_TEMP_SETPOINT"], T_ROOM, K["CONDUCTION_COEFF"], 1.0e3, dt) / dt
    assert close(p, EX["T7_CONDUCTION_20C"], EX["T7_CONDUCTION_20C_TOL"])
    assert close(K["CELL_ENERGY_MAX"] / p, EX["T7_ENDURANCE_20C"], EX["T7_ENDURANCE_20C_TOL"])
    # a tiny voxel cannot take more than equalises it: e -> C (Tc - Tg)
    c = 1e-12
    assert close(le(380.0, 300.0, K["CONDUCTION_COEFF"], c, 10.0), c * 80.0, 1e-12)
    assert le(300.0, 300.0, K["CONDUCTION_COEFF"], 1.0, dt) == 0.0


def test_t26_taxis_lag_and_tumble(lib):
    te = sig(lib, "taxis_ema_update", F, [F, D, D])
    for dt in (1.0e-3, 1.0e-2, 1.0e-1):
        ema = 0.0
        for _ in range(int(K["TAXIS_MEMORY_TIME"] / dt + 0.5)):
            ema = te(ema, 1.0, dt)
        assert close(ema, EX["T26_EMA_STEP_RESPONSE"], 1e-5)
    cr = sig(lib, "cell_rng_init", U64, [U64, U64])
    ci = sig(lib, "cell_stream_inc", U64, [U64])
    tt = sig(lib, "taxis_tumble_angle", D, [PU64, U64])
    st, inc = U64(cr(20260802, 7)), ci(7)
    n = 200000
    angles = np.array([tt(C.byref(st), inc) for _ in range(n)])
    assert angles.max() <= math.pi and angles.min() >= 0.0
    assert close(float(angles.mean()), EX["T26_TUMBLE_MEAN_CLAMPED"], EX["T26_TUMBLE_MEAN_CLAMPED_TOL"])


def test_t18_saturated_uptake_through_the_lifecycle(oracle):
    # One live cell in CO2 a million times the half-saturation: its store grows
    # by the saturated rate times the biology step (test_lifecycle.cu:94).
    conc = 1.0e6 * K["LIFE_CO2_HALF_SATURATION"]
    with ah.World(oracle, capacity=4, co2_init=conc, thermal_enabled=0, emission_enabled=0, taxis_enabled=0,
                  contact_enabled=0, thermal_noise=0,
                  grid={ah.TEMPERATURE: (8, 8, 1), ah.CO2: (8, 8, 1), ah.N2: (-1, 0, 0),
                        ah.IRRADIANCE: (8, 8, 1)}) as w:
        w.spawn(1, 3, charge_dist=ah.DIST_CONSTANT, charge_a=0.5)
        w.step(1)
        held = float(w.download()["co2_held"][0])
        dt_bio = w.dt * w.desc.biology_rate
    rate = held / dt_bio
    assert close(rate, EX["T18_UPTAKE_MAX"], 2e-6)
    assert close(K["LIFE_CO2_UPTAKE_MAX"], EX["T18_UPTAKE_MAX"], EX["T18_UPTAKE_MAX_TOL"])


def _gauss_world(oracle, n=64, sigma_cells=5.0):
    w = ah.World(oracle, capacity=4, co2_init=0.0,
                 grid={ah.TEMPERATURE: (16, 16, 1), ah.CO2: (n, n, n), ah.N2: (-1, 0, 0), ah.IRRADIANCE: (8, 8, 1)},
                 chamber_w=550e-6, chamber_h=550e-6, chamber_d=550e-6)
    (nx, ny, nz), (dx, dy, dz) = w.field_dims(ah.CO2)
    c = (np.arange(n) + 0.5 - 0.5 * n) * dx
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    s0 = sigma_cells * dx
    u = 100.0 * np.exp(-(xx ** 2 + yy ** 2 + zz ** 2) / (2 * s0 ** 2))
    w.set_field(ah.CO2, u.astype(np.float32))
    return w, c, s0


def test_ftcs_gaussian_spreading_and_conservation(oracle):
    # test_fields.cu:119-156 in 3D: sigma^2 grows by 2 D t per axis, and the
    # insulated (Neumann) CO2 field conserves its total.
    w, c, s0 = _gauss_world(oracle)
    with w:
        u0 = w.field(ah.CO2).astype(np.float64)
        dt = 1e-3
        t_run = 0.3
        for _ in range(int(round(t_run / dt))):
            w.stage("diffuse", ah.CO2, dt)
        u1 = w.field(ah.CO2).astype(np.float64)
    assert close(u1.sum(), u0.sum(), 1e-5)
    assert u1.min() >= -1e-4 and u1.max() <= 100.0 + 1e-3
    for axis in range(3):
        prof = u1.sum(axis=tuple(a for a in range(3) if a != axis))
        var = float((prof * c ** 2).sum() / prof.sum())
        want = s0 ** 2 + 2.0 * K["CO2_DIFFUSIVITY_WATER"] * t_run
        assert close(var, want, 2e-2), (axis, var, want)
