"""Independent pins of the oracle's composed stages.

The GPU parity tests compare the HIP library with the oracle, which by itself
says nothing about whether the oracle follows the reference. The golden tests
(test_oracle_golden.py) pin the oracle's building blocks against the
reference's own compiled headers and physics expectations. This module pins
the stages those blocks are composed into, each against a restatement that
shares no code with the oracle:

  * contact forces: an O(n^2) pair sum (tests/physics/test_contact.cu:23-68,
    test_hash.cu:83-128: every neighbour inside the contact range, once);
  * the spatial hash: numpy's stable argsort of bucket keys computed here
    (src/sim/hash.cu:15-38, hash.cuh:63-78; CUB's stable sort is uniquely
    defined);
  * division and compaction: a numpy exclusive scan and stable filter
    (src/sim/lifecycle.cu:119-204, :269-293; src/sim/cell_store.cu:368-456);
  * the FTCS boundary ghosts: closed forms for a uniform field after one and
    two substeps, and Newton cooling of the insulated-but-convective chamber
    (src/fields/diffuse.cu:19-44, :91-119; test_fields.cu:158-189);
  * field sampling: trilinear and bilinear interpolation reproduce an affine
    field exactly, and clamp at the edges (src/fields/grid.cuh:109-125);
  * the light sweep: the Beer-Lambert product over columns of known occlusion,
    exactly 0 behind a full block, in all four directions
    (src/sim/emission.cu:17-53, :145-176), and the population depth gradient
    (test_emission.cu:161-216);
  * spawn: the draw order of spawn_kernel (src/sim/cell_store.cu:42-130)
    replayed in Python over the PCG32 primitives, which ref_vectors.json pins;
  * lifecycle at population scale: one doubling per LIFE_DOUBLING_TIME, and
    CO2 never negative when rationed (test_lifecycle.cu:114-191);
  * the thermostat: the ignition latch (test_thermal.cu:136-163) and the
    medium pinned at the setpoint (:165-205);
  * the stats ledger: order independence under any slot permutation and
    equality with a fixed-point sum computed here (test_stats.cu:90-116,
    src/sim/stats.cu:27-60, :102-153).
A dropped term, a wrong sign, a transposed axis or an off-by-one in any of
these stages fails at least one test.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from conftest import ah
from test_oracle_golden import EX, K, T_ROOM

OCC, ALIVE, AWAKE = ah.FLAG_OCCUPIED, ah.FLAG_ALIVE, ah.FLAG_AWAKE
SMALL = {ah.TEMPERATURE: (16, 16, 1), ah.CO2: (8, 8, 1), ah.N2: (-1, 0, 0), ah.IRRADIANCE: (16, 16, 1)}


def prim(oracle):
    """The oracle's PCG32 / splitmix / division-axis primitives, each pinned
    bit for bit against the reference's headers by ref_vectors.json."""
    lib = oracle.lib
    U64, PU64, PD = C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_double)
    for name, res, args in (("pcg32_next", C.c_uint32, [PU64, U64]), ("cell_stream_inc", U64, [U64]),
                            ("cell_rng_init", U64, [U64, U64]), ("pcg_split", U64, [U64, U64]),
                            ("division_axis", None, [U64, PD, PD, PD])):
        f = getattr(lib, "oracle_" + name)
        f.restype, f.argtypes = res, args
    return lib


# ---------------------------------------------------------------------------
# Contact forces: brute-force pair sum
# ---------------------------------------------------------------------------
def brute_contact(c: dict, k: float) -> np.ndarray:
    """Sum over every occupied pair closer than a diameter of
    k (2a - r) (p_i - p_j) / r (src/sim/integrator.cuh:210-219)."""
    two_a = K["CELL_DIAMETER"]
    p = np.stack([c["x"], c["y"], c["z"]], axis=1)
    occ = (c["flags"] & OCC) != 0
    idx = np.nonzero(occ)[0]
    out = np.zeros_like(p)
    for i in idx:
        d = p[i] - p[idx]
        r2 = (d * d).sum(axis=1)
        m = (r2 < two_a * two_a) & (r2 > 0.0)
        r = np.sqrt(r2[m])
        out[i] = (d[m] * (k * (two_a - r) / r)[:, None]).sum(axis=0)
    return out


@pytest.mark.parametrize("hash_cell", [0.0, 1e-5])
def test_contact_forces_equal_the_pair_sum(oracle, hash_cell):
    # A dense cluster (many overlaps), with free slots and corpses: corpses
    # push and are pushed, free slots neither (integrator.cu:47-66).
    kw = dict(capacity=2500, chamber_w=1e-3, chamber_h=1e-3, hash_cell_size=hash_cell, grid=SMALL,
              physics_rate=0.5)
    with ah.World(oracle, **kw) as w:
        w.spawn(2500, 7, placement=ah.PLACE_GAUSSIAN, place_radius=60e-6)
        cells = w.download()
    cells["flags"][::37] = 0
    cells["flags"][5::11] &= ~np.uint32(ALIVE)
    forces = {}
    for contact in (1, 0):
        with ah.World(oracle, contact_enabled=contact, **kw) as w:
            w.upload(cells)
            w.stage("forces")
            forces[contact] = np.stack(w.forces(), axis=1)
    got = forces[1] - forces[0]
    want = brute_contact(cells, K["CONTACT_STIFFNESS"] / 0.5)
    touching = np.count_nonzero(np.abs(want).sum(axis=1))
    assert touching > 1500, touching
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-9 * scale
    # free slots carry no force at all
    assert not np.any(forces[1][cells["flags"] == 0])


# ---------------------------------------------------------------------------
# Spatial hash: numpy's stable sort of keys computed here
# ---------------------------------------------------------------------------
def hash_keys(c: dict, nx: int, ny: int, nz: int, cell: float) -> np.ndarray:
    """bucket = clamp(int32((p - origin) * inv)), origin and inv rounded to
    float as HashView carries them (src/sim/hash.cu:95-103, hash.cuh:63-78);
    free slots get the sentinel nb (hash.cu:15-25)."""
    inv = float(np.float32(1.0 / cell))
    key = np.zeros(len(c["x"]), np.int64)
    for axis, n, stride in (("x", nx, 1), ("y", ny, nx), ("z", nz, nx * ny)):
        origin = float(np.float32(-0.5 * n * cell))
        g = np.trunc((c[axis] - origin) * inv)
        key += np.clip(g, 0, n - 1).astype(np.int64) * stride
    key[(c["flags"] & OCC) == 0] = nx * ny * nz
    return key


@pytest.mark.parametrize("hash_cell", [0.0, 1e-5])
def test_hash_is_the_stable_sort_of_bucket_keys(oracle, hash_cell):
    rng = np.random.default_rng(5)
    n = 6000
    kw = dict(capacity=n, hash_cell_size=hash_cell, grid=SMALL)
    with ah.World(oracle, **kw) as w:
        nx, ny, nz, cell = w.hash_dims()
        d = w.desc
        # clusters (shared buckets), cells past the grid (clamped) and free slots
        x = np.concatenate([rng.uniform(-0.6, 0.6, n // 2) * d.chamber_w, rng.normal(0.0, 3e-5, n - n // 2)])
        y = np.concatenate([rng.uniform(-0.6, 0.6, n // 2) * d.chamber_h, rng.normal(1e-4, 3e-5, n - n // 2)])
        z = rng.uniform(-0.6, 0.6, n) * d.chamber_d
        flags = np.full(n, OCC | ALIVE, np.uint32)
        flags[rng.random(n) < 0.03] = 0
        flags[rng.random(n) < 0.05] &= ~np.uint32(ALIVE)
        cells = {"id": np.arange(1, n + 1, dtype=np.uint64), "flags": flags, "x": x, "y": y, "z": z}
        w.upload(cells)
        w.hash_build()
        vals, bs, be = w.hash_download()
    keys = hash_keys(cells, nx, ny, nz, cell)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(vals, order.astype(np.uint32))
    sk = keys[order]
    nb = nx * ny * nz
    want_s = np.zeros(nb, np.int32)
    want_e = np.zeros(nb, np.int32)
    occupied = np.unique(sk[sk < nb])
    want_s[occupied] = np.searchsorted(sk, occupied, side="left")
    want_e[occupied] = np.searchsorted(sk, occupied, side="right")
    assert np.array_equal(bs, want_s) and np.array_equal(be, want_e)
    assert len(occupied) > 500


def test_neighbour_walk_finds_every_contact_once(oracle):
    """test_hash.cu:83-128: within the 27 buckets around a cell, every cell
    closer than a bucket is present exactly once, ascending slot per bucket."""
    rng = np.random.default_rng(9)
    n = 3000
    with ah.World(oracle, capacity=n, chamber_w=6e-4, chamber_h=6e-4, grid=SMALL) as w:
        w.spawn(n, 3, placement=ah.PLACE_GAUSSIAN, place_radius=8e-5)
        c = w.download()
        w.hash_build()
        vals, bs, be = w.hash_download()
        nx, ny, nz, cell = w.hash_dims()
    keys = hash_keys(c, nx, ny, nz, cell)
    p = np.stack([c["x"], c["y"], c["z"]], axis=1)
    for i in rng.choice(n, 200, replace=False):
        gx, gy, gz = keys[i] % nx, (keys[i] // nx) % ny, keys[i] // (nx * ny)
        seen = []
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    X, Y, Z = gx + dx, gy + dy, gz + dz
                    if 0 <= X < nx and 0 <= Y < ny and 0 <= Z < nz:
                        b = (Z * ny + Y) * nx + X
                        run = vals[bs[b]:be[b]]
                        assert np.all(np.diff(run.astype(np.int64)) > 0)
                        seen.extend(run.tolist())
        assert len(seen) == len(set(seen))
        r = np.sqrt(((p - p[i]) ** 2).sum(axis=1))
        near = set(np.nonzero(r < cell)[0].tolist())
        assert near <= set(seen), sorted(near - set(seen))[:5]


# ---------------------------------------------------------------------------
# Division and compaction: numpy exclusive scan and stable filter
# ---------------------------------------------------------------------------
def divide_and_compact(oracle, c: dict, n: int, cap: int, next_id: int, offset: float, chamber, compaction: bool):
    lib = prim(oracle)
    quota = K["CO2_MASS_PER_DIVISION"]
    a_clamp = K["CELL_RADIUS"]
    out = {k: np.concatenate([v[:n], np.zeros(cap - n, v.dtype)]) for k, v in c.items()}
    f = out["flags"][:n]
    occ, alive = (f & OCC) != 0, (f & ALIVE) != 0
    div = occ & alive & (out["co2_held"][:n] >= quota)
    prefix = np.cumsum(div) - div
    births = min(int(div.sum()), cap - n)
    deads = int((occ & ~alive).sum())
    hx, hy, hz = (0.5 * s - a_clamp for s in chamber)
    cl = lambda v, h: -h if v < -h else (h if v > h else v)  # noqa: E731
    cost = K["LIFE_DIVISION_ENERGY_COST"]
    for i in np.nonzero(div)[0]:
        p = int(prefix[i])
        if p >= births:
            continue
        s = n + p
        did = next_id + p
        half = 0.5 * (out["energy"][i] - cost)
        half = half if half > 0.0 else 0.0
        out["energy"][i] = out["energy"][s] = half
        out["co2_held"][i] = out["co2_held"][s] = 0.0
        out["biomass"][s] = K["CELL_MASS_DRY"]
        ax, ay, az = C.c_double(), C.c_double(), C.c_double()
        lib.oracle_division_axis(did, C.byref(ax), C.byref(ay), C.byref(az))
        for key, u, h in (("x", ax.value, hx), ("y", ay.value, hy), ("z", az.value, hz)):
            parent = out[key][i]
            out[key][s] = cl(parent + offset * u, h)
            out[key][i] = cl(parent - offset * u, h)
        for key in ("vx", "vy", "vz", "flags", "temp_cell", "dir_x", "dir_y", "dir_z", "t_local", "co2_local",
                    "n2_local"):
            out[key][s] = out[key][i]
        out["id"][s] = did
        out["death_cause"][s] = 0
        for key in ("emit_power", "age_s", "taxis_memory", "run_timer", "irradiance"):
            out[key][s] = 0
        out["rng_state"][s] = lib.oracle_pcg_split(int(out["rng_state"][i]), did)
    m = n + births
    out = {k: v[:m] for k, v in out.items()}
    if compaction and deads > 0:
        keep = ((out["flags"] & OCC) != 0) & ((out["flags"] & ALIVE) != 0)
        out = {k: v[keep] for k, v in out.items()}
    return out, births, deads


@pytest.mark.parametrize("cap_extra,compaction,offset", [(200, 1, 0.0), (200, 0, 0.0), (37, 1, 0.5e-6)])
def test_division_and_compaction_match_scan_and_filter(oracle, cap_extra, compaction, offset):
    n = 500
    cap = n + cap_extra
    chamber = (1e-3, 1e-3, K["CHAMBER_D"])
    kw = dict(capacity=cap, chamber_w=chamber[0], chamber_h=chamber[1], co2_init=0.0, compaction_enabled=compaction,
              division_offset=offset, grid=SMALL, biology_rate=1e6)
    rng = np.random.default_rng(11 + cap_extra)
    with ah.World(oracle, **kw) as w:
        w.spawn(n, 4, charge_dist=ah.DIST_UNIFORM, charge_b=0.9)
        c = w.download()
        quota = K["CO2_MASS_PER_DIVISION"]
        c["co2_held"] = rng.uniform(0.0, 0.99, n) * quota
        div = rng.random(n) < 0.3
        c["co2_held"][div] = quota * rng.uniform(1.0, 1.5, div.sum())
        c["co2_held"][div.nonzero()[0][:5]] = quota                      # exactly at the quota divides
        c["flags"][::50] = 0                                              # free slots
        c["flags"][3::13] &= ~np.uint32(ALIVE)                            # corpses
        c["t_local"][:] = 300.0                                           # nobody overheats
        hw = 0.5 * chamber[0] - 0.2 * K["CELL_RADIUS"]                    # parents at the walls: clamps
        c["x"][div.nonzero()[0][5:25]] = hw
        c["y"][div.nonzero()[0][25:45]] = -hw
        w.upload(c)
        ctr = w.counters()
        ctr.next_id = 10_000
        w.set_counters(ctr)
        w.stage("lifecycle", 1e-3)
        got = w.download()
        ctr = w.counters()
    want, births, deads = divide_and_compact(oracle, c, n, cap, 10_000, offset or K["CELL_RADIUS"], chamber,
                                             bool(compaction))
    assert births > 30 and deads > 30
    if cap_extra == 37:
        assert births == 37                                              # the store filled: births clamped
    assert ctr.count == len(want["id"])
    assert ctr.next_id == 10_000 + births
    assert ctr.divisions_total == births and ctr.deaths_total == deads
    for key in want:
        assert np.array_equal(got[key], want[key]), key


# ---------------------------------------------------------------------------
# FTCS boundaries: closed forms on a uniform field
# ---------------------------------------------------------------------------
def robin(d):
    return d * K["AIR_CONVECTION_H"] / K["WATER_CONDUCTIVITY"]


def face_counts(n):
    f = np.zeros(n)
    f[0] += 1
    f[-1] += 1
    return f


@pytest.mark.parametrize("dims", [(64, 64, 1), (48, 40, 24)])
def test_ftcs_one_substep_on_a_uniform_field(oracle, dims):
    """A uniform field u0 = ambient + e after one substep: interior voxels
    keep u0 exactly; a voxel on a Robin (temperature) wall loses
    c_axis * robin_axis * e per wall it touches (ghost = u - robin (u -
    ambient)); every voxel of an insulated (Neumann, CO2) field keeps u0."""
    amb, e = 300.0, 200.0
    chamber = (550e-6, 500e-6, 200e-6) if dims[2] > 1 else (1e-3, 1e-3, K["CHAMBER_D"])
    grid = {ah.TEMPERATURE: dims, ah.CO2: dims, ah.N2: (-1, 0, 0), ah.IRRADIANCE: (8, 8, 1)}
    with ah.World(oracle, capacity=4, ambient_temp=amb, chamber_w=chamber[0], chamber_h=chamber[1],
                  chamber_d=chamber[2], grid=grid) as w:
        (nx, ny, nz), (dx, dy, dz) = w.field_dims(ah.TEMPERATURE)
        D = K["WATER_THERMAL_DIFFUSIVITY"]
        inv = 1 / dx ** 2 + 1 / dy ** 2 + (1 / dz ** 2 if nz > 1 else 0.0)
        dt_max = dx * dx / (4 * D) if nz == 1 else 1.0 / (2 * D * inv)
        dt = 0.9 * dt_max                                                 # one substep
        u0 = np.float32(amb + e)
        w.set_field(ah.TEMPERATURE, np.full(nx * ny * nz, u0, np.float32))
        w.set_field(ah.CO2, np.full(nx * ny * nz, u0, np.float32))
        w.stage("diffuse", ah.TEMPERATURE, dt)
        w.stage("diffuse", ah.CO2, dt)
        T, Cf = w.field(ah.TEMPERATURE).astype(np.float64), w.field(ah.CO2)
    assert np.all(Cf == u0)
    loss = (D * dt / dx ** 2 * robin(dx) * face_counts(nx)[None, None, :]
            + D * dt / dy ** 2 * robin(dy) * face_counts(ny)[None, :, None])
    if nz > 1:
        loss = loss + D * dt / dz ** 2 * robin(dz) * face_counts(nz)[:, None, None]
    want = float(u0) - e * loss
    assert np.abs(T - want).max() <= 1.2e-4
    assert T[nz // 2, ny // 2, nx // 2] == u0
    assert (float(u0) - T).max() > 40 * 1.2e-4                            # the walls really lose heat


def test_ftcs_substep_rule_and_two_substeps(oracle):
    """3D substep count: S = floor(dt / dt_max) + 1 with dt_max = 1 / (2 D
    sum 1/dx_i^2). At dt = 1.5 dt_max it is 2, and a uniform field then
    follows the two-substep closed form next to an x wall (voxels away from
    the y and z walls): f1 = u0 - c r e, f2 = f1 + c (g2 + u0 - 2 f1) with
    g2 = f1 - r (f1 - ambient), and the next voxel in u0 - c^2 r e."""
    amb, e = 300.0, 200.0
    dims = (48, 40, 24)
    grid = {ah.TEMPERATURE: dims, ah.CO2: (8, 8, 1), ah.N2: (-1, 0, 0), ah.IRRADIANCE: (8, 8, 1)}
    with ah.World(oracle, capacity=4, ambient_temp=amb, chamber_w=550e-6, chamber_h=500e-6, chamber_d=200e-6,
                  grid=grid) as w:
        (nx, ny, nz), (dx, dy, dz) = w.field_dims(ah.TEMPERATURE)
        D = K["WATER_THERMAL_DIFFUSIVITY"]
        dt_max = 1.0 / (2 * D * (1 / dx ** 2 + 1 / dy ** 2 + 1 / dz ** 2))
        dt = 1.5 * dt_max
        u0 = amb + e
        w.set_field(ah.TEMPERATURE, np.full(nx * ny * nz, u0, np.float32))
        w.stage("diffuse", ah.TEMPERATURE, dt)
        T = w.field(ah.TEMPERATURE).astype(np.float64)
    c = D * (dt / 2) / dx ** 2
    r = robin(dx)
    f1 = u0 - c * r * e
    g2 = f1 - r * (f1 - amb)
    f2 = f1 + c * (g2 + u0 - 2 * f1)
    inner = (slice(2, nz - 2), slice(2, ny - 2))
    assert np.abs(T[inner + (0,)] - f2).max() <= 1.2e-4
    assert np.abs(T[inner + (nx - 1,)] - f2).max() <= 1.2e-4
    assert np.abs(T[inner + (1,)] - (u0 - c * c * r * e)).max() <= 1.2e-4
    assert np.all(T[inner + (slice(3, nx - 3),)] == np.float32(u0))


def test_ftcs_convective_walls_cool_as_newton_predicts(oracle):
    """test_fields.cu:158-189 with closed forms: an insulated (Neumann)
    field keeps its mean; behind the Robin walls the mean decays toward
    ambient as a lumped body, d(mean)/dt = -(4 h / (rho c_p W)) (mean -
    ambient) for an n x n grid of side W (the flux through each wall face
    is D r (u - ambient) / dx^2 per unit of c, and D / k = 1 / (rho c_p))."""
    amb, e = 300.0, 50.0
    grid = {ah.TEMPERATURE: (128, 128, 1), ah.CO2: (128, 128, 1), ah.N2: (-1, 0, 0), ah.IRRADIANCE: (8, 8, 1)}
    with ah.World(oracle, capacity=4, ambient_temp=amb, grid=grid) as w:
        n = 128 * 128
        w.set_field(ah.TEMPERATURE, np.full(n, amb + e, np.float32))
        w.set_field(ah.CO2, np.full(n, amb + e, np.float32))
        steps, dt = 5000, 1e-3
        for _ in range(steps):
            w.stage("diffuse", ah.TEMPERATURE, dt)
            w.stage("diffuse", ah.CO2, dt)
        t_mean = w.field(ah.TEMPERATURE).astype(np.float64).mean()
        c_mean = w.field(ah.CO2).astype(np.float64).mean()
    assert abs(c_mean - (amb + e)) <= 1e-4
    tau = K["WATER_DENSITY"] * K["WATER_SPECIFIC_HEAT"] * K["CHAMBER_W"] / (4.0 * K["AIR_CONVECTION_H"])
    want_drop = e * (1.0 - math.exp(-steps * dt / tau))
    got_drop = amb + e - t_mean
    assert abs(got_drop - want_drop) <= 0.03 * want_drop, (got_drop, want_drop)


# ---------------------------------------------------------------------------
# Field sampling: exact on affine fields
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dims", [(40, 32, 16), (64, 64, 1)])
def test_field_sample_reproduces_affine_fields(oracle, dims):
    """T = 300 + g . (voxel index) is reproduced at any point between the
    first and last voxel centres (bi/trilinear weights), and held at the
    edge value beyond them (edge-clamped indices)."""
    nx, ny, nz = dims
    chamber = (1e-3, 0.8e-3, 2e-4) if nz > 1 else (1e-3, 1e-3, K["CHAMBER_D"])
    grid = {ah.TEMPERATURE: dims, ah.CO2: (8, 8, 1), ah.N2: (-1, 0, 0), ah.IRRADIANCE: (8, 8, 1)}
    gxyz = (0.37, 1.3, 2.9)
    n = 3000
    with ah.World(oracle, capacity=n, chamber_w=chamber[0], chamber_h=chamber[1], chamber_d=chamber[2],
                  grid=grid, contact_enabled=0) as w:
        (_, _, _), (dx, dy, dz) = w.field_dims(ah.TEMPERATURE)
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        field = 300.0 + gxyz[0] * i + gxyz[1] * j + (gxyz[2] * k if nz > 1 else 0.0)
        w.set_field(ah.TEMPERATURE, field.astype(np.float32))
        rng = np.random.default_rng(2)
        pos = [rng.uniform(-0.5, 0.5, n) * s for s in chamber]
        cells = {"id": np.arange(1, n + 1, dtype=np.uint64), "flags": np.full(n, OCC | ALIVE, np.uint32),
                 "x": pos[0], "y": pos[1], "z": pos[2], "biomass": np.full(n, K["CELL_MASS_DRY"])}
        w.upload(cells)
        w.stage("forces")
        got = w.download()["t_local"].astype(np.float64)

    def index(p, size, d, m):
        g = (p + 0.5 * size) / d - 0.5
        return np.clip(g, 0.0, m - 1)

    want = (300.0 + gxyz[0] * index(pos[0], chamber[0], dx, nx) + gxyz[1] * index(pos[1], chamber[1], dy, ny)
            + (gxyz[2] * index(pos[2], chamber[2], dz, nz) if nz > 1 else 0.0))
    assert np.abs(got - want).max() <= 3e-4
    clamped = (np.abs(pos[0]) > 0.5 * chamber[0] - 0.5 * dx) | (np.abs(pos[1]) > 0.5 * chamber[1] - 0.5 * dy)
    assert clamped.sum() > 50


# ---------------------------------------------------------------------------
# Light: Beer-Lambert columns of known occlusion
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("light", [(1.0, 0.0), (-1.0, 0.1), (0.2, 1.0), (0.0, -1.0)])
def test_light_sweep_is_the_beer_lambert_product(oracle, light):
    """Cells stamped at known bins of one line: the irradiance at step s of
    the march is source * prod over the bins before it of max(0, 1 - A/face),
    A = (cells x pi a^2, corpses x (1 - albedo_dead)), face = dx x depth; a
    bin holding a cell still sees the light that reached its near face; a
    line blocked by more than its face is exactly 0.0 behind the block; lines
    with no cells carry the source exactly."""
    n, W = 64, 1e-3
    grid = {ah.TEMPERATURE: (16, 16, 1), ah.CO2: (8, 8, 1), ah.N2: (-1, 0, 0), ah.IRRADIANCE: (n, n, 1)}
    source = 1.0e4
    axis = 0 if abs(light[0]) >= abs(light[1]) else 1
    sign = 1 if (light[axis] >= 0.0) else -1
    dx = W / n
    centre = lambda b: -0.5 * W + (b + 0.5) * dx  # noqa: E731
    # (line, bin along the light, cells, corpses); line 40 is blocked fully at bin 30
    stamps = [(10, 5, 2, 0), (10, 12, 1, 0), (10, 20, 3, 0), (10, 33, 0, 1), (10, 47, 1, 0), (10, 63, 2, 0),
              (40, 30, 14, 0), (40, 50, 1, 0), (41, 0, 1, 2)]
    xs, ys, fl = [], [], []
    for line, b, alive, dead in stamps:
        along = b if sign > 0 else n - 1 - b
        for k in range(alive + dead):
            p_along, p_line = centre(along), centre(line)
            xs.append(p_along if axis == 0 else p_line)
            ys.append(p_line if axis == 0 else p_along)
            fl.append(OCC | ALIVE if k < alive else OCC)
    m = len(xs)
    cells = {"id": np.arange(1, m + 1, dtype=np.uint64), "flags": np.array(fl, np.uint32), "x": np.array(xs),
             "y": np.array(ys), "z": np.zeros(m), "biomass": np.full(m, K["CELL_MASS_DRY"])}
    with ah.World(oracle, capacity=m, chamber_w=W, chamber_h=W, grid=grid, light_enabled=1, light_irradiance=source,
                  light_dir_x=light[0], light_dir_y=light[1], ambient_irradiance=0.0, thermal_enabled=0) as w:
        w.upload(cells)
        w.stage("emission", 1e-3)
        irr = w.field(ah.IRRADIANCE)[0]                                  # (y, x)
        depth = w.desc.chamber_d
    lines = irr if axis == 0 else irr.T                                   # (line, position)
    face = dx * depth
    a_cell = K["CELL_CROSS_SECTION"]
    for line in range(n):
        area = np.zeros(n)
        for ln, b, alive, dead in stamps:
            if ln == line:
                area[b] += alive * a_cell + dead * a_cell * (1.0 - K["CELL_ALBEDO_DEAD"])
        t, want = 1.0, np.zeros(n)
        for s in range(n):
            want[s] = np.float32(source * t)
            t *= max(0.0, 1.0 - area[s] / face)
        got = lines[line] if sign > 0 else lines[line][::-1]
        if not area.any():
            assert np.all(got == np.float32(source)), line
        else:
            assert np.allclose(got, want, rtol=2e-7, atol=0.0), (line, got[:8], want[:8])
    blocked = lines[40] if sign > 0 else lines[40][::-1]
    assert np.all(blocked[31:] == 0.0) and blocked[30] > 0.0


def test_depth_along_the_light_decides_who_charges(oracle):
    """test_emission.cu:161-216: 8,000 cells lit along +x at 5e3 W/m^2
    from zero charge, thermal, contact and adhesion off, 500 ticks: charge
    falls with depth (r < -0.5) and the lit face holds over 5x the far side."""
    n = 8000
    with ah.World(oracle, capacity=n, seed=4242, thermal_enabled=0, contact_enabled=0, adhesion_enabled=0,
                  light_enabled=1, light_irradiance=5.0e3, light_dir_x=1.0, light_dir_y=0.0) as w:
        w.spawn(n, 4242)
        c = w.download()
        c["energy"][:] = 0.0
        w.upload(c)
        w.step(500)
        c = w.download()
    x, e = c["x"], c["energy"]
    r = np.corrcoef(x, e)[0, 1]
    front = e[x < -0.3 * K["CHAMBER_W"]].mean()
    back = e[x > 0.3 * K["CHAMBER_W"]].mean()
    assert r < -0.5, r
    assert front > 5.0 * back, (front, back)


# ---------------------------------------------------------------------------
# Spawn: the draw order replayed over the pinned PCG32 primitives
# ---------------------------------------------------------------------------
class Stream:
    def __init__(self, lib, seed: int, cid: int):
        self.lib = lib
        self.state = C.c_uint64(lib.oracle_cell_rng_init(seed, cid))
        self.inc = lib.oracle_cell_stream_inc(cid)

    def u32(self) -> int:
        return self.lib.oracle_pcg32_next(C.byref(self.state), self.inc)

    def u01(self) -> float:
        hi = self.u32() << 21
        lo = self.u32() >> 11
        return float((hi ^ lo) & ((1 << 53) - 1)) * (1.0 / 9007199254740992.0)

    def uniform(self, lo, hi):
        return lo + (hi - lo) * self.u01()

    def pair(self):
        u1 = max(self.u01(), 1.0e-300)
        u2 = self.u01()
        mag = math.sqrt(-2.0 * math.log(u1))
        ang = 6.283185307179586476925286766559 * u2
        return mag * math.cos(ang), mag * math.sin(ang)


@pytest.mark.parametrize("placement,dist", [(ah.PLACE_UNIFORM, ah.DIST_UNIFORM), (ah.PLACE_GAUSSIAN, ah.DIST_NORMAL),
                                            (ah.PLACE_UNIFORM, ah.DIST_LOG_UNIFORM)])
def test_spawn_follows_the_reference_draw_order(oracle, placement, dist):
    """Per cell: stream from (seed, id); placement draws (x then y; a
    Gaussian pair for Gaussian placement), clamp, z, the charge, then two
    Gaussian pairs of which three values make the axis
    (src/sim/cell_store.cu:42-130). Positions, charge and the stream state
    after the draws are exact; the axis is normalized and rounded to float."""
    lib = prim(oracle)
    seed, n = 77, 300
    charge_a, charge_b = (0.05, 0.3) if dist != ah.DIST_NORMAL else (0.4, 0.1)
    if dist == ah.DIST_LOG_UNIFORM:
        charge_a, charge_b = 1e-3, 0.2
    with ah.World(oracle, capacity=n, grid=SMALL) as w:
        w.spawn(n, seed, placement=placement, place_x=1e-4, place_y=-2e-4, place_radius=3e-4, charge_dist=dist,
                charge_a=charge_a, charge_b=charge_b)
        got = w.download()
        d = w.desc
    a = K["CELL_RADIUS"]
    clamp = lambda v, lo, hi: lo if v < lo else (hi if v > hi else v)  # noqa: E731
    for i in range(n):
        cid = i + 1
        r = Stream(lib, seed, cid)
        if placement == ah.PLACE_GAUSSIAN:
            gx, gy = r.pair()
            x, y = 1e-4 + 3e-4 * gx, -2e-4 + 3e-4 * gy
        else:
            x = r.uniform(-0.5 * d.chamber_w, 0.5 * d.chamber_w)
            y = r.uniform(-0.5 * d.chamber_h, 0.5 * d.chamber_h)
        x = clamp(x, -0.5 * d.chamber_w + a, 0.5 * d.chamber_w - a)
        y = clamp(y, -0.5 * d.chamber_h + a, 0.5 * d.chamber_h - a)
        z = r.uniform(-0.5 * d.chamber_d + a, 0.5 * d.chamber_d - a)
        if dist == ah.DIST_UNIFORM:
            q = r.uniform(charge_a, charge_b)
        elif dist == ah.DIST_NORMAL:
            q = charge_a + charge_b * r.pair()[0]
        else:
            q = math.exp(r.uniform(math.log(charge_a), math.log(charge_b)))
        q = clamp(q, 0.0, 1.0)
        g0, g1 = r.pair()
        g2, _ = r.pair()
        length = math.sqrt(g0 * g0 + g1 * g1 + g2 * g2)
        assert (got["x"][i], got["y"][i], got["z"][i]) == (x, y, z), i
        assert got["energy"][i] == q * K["CELL_ENERGY_MAX"], i
        assert int(got["rng_state"][i]) == r.state.value, i
        for key, g in (("dir_x", g0), ("dir_y", g1), ("dir_z", g2)):
            assert abs(float(got[key][i]) - g / length) <= 6e-8, (i, key)
        assert int(got["id"][i]) == cid


# ---------------------------------------------------------------------------
# Lifecycle at population scale
# ---------------------------------------------------------------------------
def life_world(oracle, n, co2, rate=2.0e7, seed=20260802):
    """make_world of test_lifecycle.cu:22-41: room to grow, mechanics, light
    and taxis off, dormant cells of charge 0.5."""
    w = ah.World(oracle, capacity=1 << 17, seed=seed, co2_init=co2, contact_enabled=0, adhesion_enabled=0,
                 emission_enabled=0, taxis_enabled=0, biology_rate=rate)
    w.spawn(n, seed, charge_dist=ah.DIST_CONSTANT, charge_a=0.5)
    return w


def test_population_doubles_in_one_doubling_time(oracle):
    rate = 2.0e7
    ticks = int(EX["T18_DOUBLING_TIME"] / (K["DT_PHYSICS"] * rate) + 0.5)
    with life_world(oracle, 4000, 1.5, rate) as w:
        w.step(ticks)
        ratio = w.counters().count / 4000
    assert abs(ratio - 2.0) <= EX["T18_DOUBLING_TIME_TOL"] * 2.0, ratio


def test_rationed_co2_never_goes_negative(oracle):
    """T18.3: a medium holding 1.5 divisions' worth per cell stops growth
    well short of the saturated control, and rationing keeps every CO2 voxel
    non-negative."""
    rate, n0 = 2.0e7, 2000
    ticks = int(2.0 * EX["T18_DOUBLING_TIME"] / (K["DT_PHYSICS"] * rate) + 0.5)
    vol = K["CHAMBER_W"] * K["CHAMBER_H"] * K["CHAMBER_D"]
    ends = []
    lo = None
    for limited in (False, True):
        conc = 1.5 * n0 * K["CO2_MASS_PER_DIVISION"] / vol if limited else 1.5
        with life_world(oracle, n0, conc, rate) as w:
            c0 = w.stats()["co2_total_kg"]
            w.step(ticks)
            ends.append(w.counters().count)
            if limited:
                lo = float(w.field(ah.CO2).min())
                c1 = w.stats()["co2_total_kg"]
    assert ends[0] > n0
    assert ends[1] - n0 < 0.25 * (ends[0] - n0), ends
    assert c1 < c0
    assert lo >= 0.0


# ---------------------------------------------------------------------------
# Thermostat
# ---------------------------------------------------------------------------
def test_ignition_latch_is_one_way(oracle):
    """test_thermal.cu:136-163 on a 64 x 64 temperature grid: cold cells
    never wake; past the setpoint all wake; chilled again they stay awake."""
    grid = {ah.TEMPERATURE: (64, 64, 1), ah.CO2: (16, 16, 1), ah.N2: (-1, 0, 0), ah.IRRADIANCE: (16, 16, 1)}
    with ah.World(oracle, capacity=500, ambient_temp=T_ROOM, grid=grid) as w:
        w.spawn(500, 6, charge_dist=ah.DIST_CONSTANT, charge_a=0.5)
        awake = lambda: int(np.count_nonzero(w.download()["flags"] & AWAKE))  # noqa: E731
        w.step(200)
        assert awake() == 0
        w.fill(ah.TEMPERATURE, K["CELL_TEMP_SETPOINT"] + 2.0)
        w.step(50)
        assert awake() == 500
        w.fill(ah.TEMPERATURE, T_ROOM)
        w.step(500)
        assert awake() == 500


def test_medium_pins_to_the_setpoint(oracle):
    """test_thermal.cu:165-205 in a small chamber on a faster clock: awake
    cells at full charge heat the medium, which rises well above ambient but
    never above the setpoint (+0.2 K) and never boils. As in the reference's
    test the cells are sparse (0.024 per temperature voxel here, 0.008
    there): the lumped exchange brings one voxel at most to the setpoint per
    substep, so several cells sharing a voxel would overshoot it together."""
    grid = {ah.TEMPERATURE: (64, 64, 1), ah.CO2: (8, 8, 1), ah.N2: (-1, 0, 0), ah.IRRADIANCE: (8, 8, 1)}
    with ah.World(oracle, capacity=100, ambient_temp=T_ROOM, chamber_w=5e-4, chamber_h=5e-4, contact_enabled=0,
                  adhesion_enabled=0, grid=grid, physics_rate=10.0) as w:
        w.spawn(100, 5, charge_dist=ah.DIST_CONSTANT, charge_a=1.0, awake=True)
        worst = -1e30
        for _ in range(20):
            w.step(20)
            s = w.stats()
            worst = max(worst, s["max_temp_medium_k"])
        assert s["mean_temp_medium_k"] > T_ROOM + 60.0
    assert worst <= K["CELL_TEMP_SETPOINT"] + 0.2, worst
    assert s["boil_events"] == 0


# ---------------------------------------------------------------------------
# Stats: order independence and the fixed-point ledger
# ---------------------------------------------------------------------------
def fixed_sum(v, scale):
    # llround(v * scale) summed as int64 (src/core/fixed_atomic.cuh:27-44)
    return int(np.sum(np.array([math.floor(abs(x) * scale + 0.5) * (1 if x >= 0 else -1) for x in v],
                               dtype=np.int64)))


def test_stats_ledger_is_order_independent(oracle):
    grid = {ah.TEMPERATURE: (64, 64, 1), ah.CO2: (32, 32, 1), ah.N2: (16, 16, 1), ah.IRRADIANCE: (16, 16, 1)}
    with ah.World(oracle, capacity=5000, grid=grid, co2_init=3e-3) as w:
        w.spawn(4000, 8, charge_dist=ah.DIST_UNIFORM, charge_b=0.9, awake=True)
        w.spawn(1000, 9, charge_dist=ah.DIST_UNIFORM, charge_b=0.9)
        w.step(20)
        c = w.download()
        fields = {f: w.field(f) for f in (ah.TEMPERATURE, ah.CO2, ah.N2)}
        s0 = w.stats()
    c["flags"][::17] &= ~np.uint32(ALIVE)
    perm = np.random.default_rng(1).permutation(len(c["id"]))
    stats = []
    for order in (np.arange(len(perm)), perm):
        with ah.World(oracle, capacity=5000, grid=grid, co2_init=3e-3) as w:
            w.upload({k: v[order] for k, v in c.items()})
            for f, v in fields.items():
                w.set_field(f, v)
            stats.append(w.stats())
    a, b = stats
    for k in ("n_live", "n_dead", "n_awake", "n_dormant", "total_energy_j", "mean_temp_cell_k",
              "mean_temp_medium_k", "max_temp_medium_k", "co2_total_kg", "n2_total_kg"):
        assert a[k] == b[k], k
    scale = K["DEPOSIT_SCALE_STATS"]
    occ = (c["flags"] & OCC) != 0
    alive = occ & ((c["flags"] & ALIVE) != 0)
    assert a["total_energy_j"] == fixed_sum(c["energy"][occ], scale) / scale
    assert a["mean_temp_cell_k"] == fixed_sum(c["temp_cell"][alive].astype(np.float64), scale) / scale / alive.sum()
    t = fields[ah.TEMPERATURE].astype(np.float64).ravel()
    assert a["mean_temp_medium_k"] == fixed_sum(t, scale) / scale / t.size
    assert s0["n_live"] == 5000 and a["n_dead"] == int(np.count_nonzero(occ & ~alive))
