"""The ASPH snapshot format on the CPU: the test-side reader (tests/asph.py)
against the reference's own contract struct and published FNV-1a vectors,
and the state hash's body order against that reader on oracle worlds.

The GPU tests (test_gpu_snapshot.py) then read the library's snapshots with
the same reader."""
import ctypes as C

import numpy as np
import pytest

import asph
from conftest import ah

# 2D reference-shaped grids: each header size is the side n.
REF2D = dict(grid={0: (128, 128, 1), 1: (64, 64, 1), 2: (32, 32, 1), 3: (128, 128, 1)})


def test_fnv1a64_published_vectors():
    # FNV-1a 64 test vectors (Fowler/Noll/Vo reference suite), from the
    # published offset basis
    basis = asph.FNV_PUBLISHED_OFFSET
    assert asph.fnv1a64(b"", basis) == 0xCBF29CE484222325
    assert asph.fnv1a64(b"a", basis) == 0xAF63DC4C8601EC8C
    assert asph.fnv1a64(b"foobar", basis) == 0x85944171F73967E8
    # chaining equals hashing the concatenation (tests/physics/test_contracts.cpp:58-61)
    assert asph.fnv1a64(b"phage", asph.fnv1a64(b"astro")) == asph.fnv1a64(b"astrophage")


def test_fnv1a64_matches_oracle(oracle):
    fn = oracle.lib.oracle_fnv1a64
    fn.restype, fn.argtypes = C.c_uint64, [C.c_void_p, C.c_size_t, C.c_uint64]
    data = np.random.default_rng(4).integers(0, 256, 4099, dtype=np.uint8).tobytes()
    assert fn(data, len(data), asph.FNV_OFFSET) == asph.fnv1a64(data)


def test_header_layout_matches_reference_struct(ref):
    fn = ref.ref_snapshot_layout
    fn.restype, fn.argtypes = C.c_int32, [C.POINTER(C.c_int64), C.c_int32]
    out = (C.c_int64 * 64)()
    n = fn(out, 64)
    v = list(out[:n])
    names = ["magic", "version", "seed", "tick", "sim_time_s", "cell_count", "cell_capacity", "taumoeba_count",
             "taumoeba_capacity", "n_temp", "n_co2", "n_n2", "chamber_w", "chamber_h", "chamber_d",
             "physics_rate", "biology_rate", "next_cell_id", "override_count", "non_canon_run", "scenario_id",
             "build_describe", "state_hash"]
    assert n == 2 + len(names) + 3
    assert v[0] == asph.HEADER_SIZE == 256
    assert v[1] == asph.OVERRIDE_SIZE == 56
    off = asph.header_offsets()
    assert {k: off[k] for k in names} == dict(zip(names, v[2:2 + len(names)]))
    assert (v[-3], v[-2]) == (asph.MAGIC, asph.VERSION)
    # the contract hash's starting value: its hash of no bytes
    assert v[-1] & 0xFFFFFFFFFFFFFFFF == asph.FNV_OFFSET


def world_body(w: ah.World, overrides: dict) -> bytes:
    fields = []
    for f in (ah.TEMPERATURE, ah.CO2, ah.N2):
        if w.field_dims(f)[0][0] > 0:
            fields.append(w.field(f).ravel())
    return asph.body_of(overrides, w.download(), ah.CELL_FIELDS, fields)


@pytest.mark.parametrize("grids", ["2d", "3d"])
def test_oracle_state_hash_is_the_body_hash(oracle, grids):
    """state_hash = FNV-1a 64 of [overrides][cell SoA in contract order]
    [T, CO2, N2] (src/sim/snapshot.cu:100-129), with a non-canon override."""
    grid = REF2D["grid"] if grids == "2d" else {0: (24, 20, 8), 1: (12, 10, 4), 2: (-1, 0, 0), 3: (24, 20, 8)}
    kw = dict(grid=grid, capacity=600, co2_mass_per_division=0.5 * 2.1e-14, co2_init=1e-3, biology_rate=1e6)
    with ah.World(oracle, **kw) as w:
        w.spawn(400, 9, charge_dist=ah.DIST_UNIFORM, charge_b=0.9, awake=True)
        w.step(3)
        body = world_body(w, {"CO2_MASS_PER_DIVISION": 0.5 * 2.1e-14})
        assert w.state_hash() == asph.fnv1a64(body)
        # canon parameters add no override record
        w2 = ah.World(oracle, **dict(kw, co2_mass_per_division=2.0999999999999999e-14))
        w2.upload(w.download())
        for f in (ah.TEMPERATURE, ah.CO2, ah.N2):
            if w.field_dims(f)[0][0] > 0:
                w2.set_field(f, w.field(f))
        assert w2.state_hash() == asph.fnv1a64(world_body(w2, {}))
        w2.close()


def test_oracle_state_hash_carries_every_override_in_order(oracle):
    """All three curated overrides non-canon: the body carries their records
    in the reference's key order, max power, flash power, CO2 per division
    (src/sim/snapshot.cu:79-83)."""
    kw = dict(REF2D, capacity=300, petrova_max_power=2.5e4, petrova_flash_power=1.0e6,
              co2_mass_per_division=0.75 * 2.1e-14)
    with ah.World(oracle, **kw) as w:
        w.spawn(200, 3, charge_dist=ah.DIST_UNIFORM, charge_b=0.5, awake=True)
        w.step(2)
        body = world_body(w, {"PETROVA_MAX_POWER": 2.5e4, "PETROVA_FLASH_POWER": 1.0e6,
                              "CO2_MASS_PER_DIVISION": 0.75 * 2.1e-14})
        assert w.state_hash() == asph.fnv1a64(body)
        assert w.stats()["non_canon_run"] == 1
    with ah.World(oracle, **dict(REF2D, capacity=10, petrova_flash_power=1.0e6)) as w:
        assert w.state_hash() == asph.fnv1a64(world_body(w, {"PETROVA_FLASH_POWER": 1.0e6}))
        assert w.stats()["non_canon_run"] == 1


def test_oracle_flash_runs_at_the_world_flash_power(oracle):
    """flash_step discharges at World::petrova_flash_power (src/sim/flash.cu:
    42, :66): a tenth of canon takes a tenth of the energy per tick."""
    canon = 3.0e6
    taken = {}
    for power in (canon, 0.1 * canon):
        with ah.World(oracle, **dict(REF2D, capacity=50, petrova_flash_power=power, emission_enabled=0,
                                          thermal_enabled=0)) as w:
            w.spawn(50, 5, charge_dist=ah.DIST_CONSTANT, charge_a=1.0, awake=True)
            e0 = w.download()["energy"].copy()
            w.set_flash(True)
            w.step(1)
            taken[power] = e0 - w.download()["energy"]
    dt = 1e-3
    assert np.allclose(taken[canon], canon * dt, rtol=1e-12)
    assert np.allclose(taken[0.1 * canon], 0.1 * canon * dt, rtol=1e-12)
