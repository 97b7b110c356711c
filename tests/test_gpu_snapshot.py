"""ASPH snapshots of GPU worlds (contracts/snapshot_v1.h, src/sim/snapshot.cu),
following the reference's T21 gates (tests/physics/test_snapshot.cu:57-181):
round-trip fidelity, bit-identical continuation across a save/restore, a
corrupt file rejected, and the in-memory form equal to the file.

The bytes are read with the test-side reader tests/asph.py, which is pinned
against the reference's own header struct and FNV in
test_snapshot_format.py: every header scalar, override record, cell array and
field value is checked against what the world reports through the ABI.
"""
import struct

import numpy as np
import pytest

import asph
from conftest import ah

pytestmark = pytest.mark.gpu

CANON_CO2_PER_DIVISION = 2.0999999999999999e-14


def growth_world(gpu, n, precision, absorbing, seed=20260802, **extra):
    """A growth world like the reference's (test_snapshot.cu:21-45):
    saturating CO2 and a compressed biology clock, so cells divide within
    tens of ticks; absorbing walls turn some into corpses that stay in
    [0, count) (compaction off)."""
    kw = dict(capacity=1 << 20, seed=seed, co2_init=1.5, emission_enabled=0, taxis_enabled=0,
              biology_rate=2.0e7, precision=precision)
    if absorbing:
        kw.update(boundary_x=ah.BOUNDARY_ABSORBING, boundary_y=ah.BOUNDARY_ABSORBING)
    kw.update(extra)
    w = ah.World(gpu, **kw)
    w.spawn(n, seed, charge_dist=ah.DIST_CONSTANT, charge_a=0.5, awake=False)
    return w, kw


def world_fields(w):
    return [w.field(f).ravel() for f in (ah.TEMPERATURE, ah.CO2, ah.N2) if w.field_dims(f)[0][0] > 0]


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_round_trip_fidelity(gpu, tmp_path, precision):
    """T21.1: a divided, partly dead population with a tuned override; every
    header scalar and the full-state hash survive save -> load."""
    w, kw = growth_world(gpu, 3000, precision, absorbing=True,
                         co2_mass_per_division=0.5 * CANON_CO2_PER_DIVISION)
    # corpses besides those the walls make: dead slots stay in [0, count)
    cells = w.download()
    dead = np.random.default_rng(1).random(len(cells["id"])) < 0.03
    cells["flags"][dead] &= ~np.uint32(ah.FLAG_ALIVE)
    w.upload(cells)
    w.step(50)
    c0 = w.counters()
    h0 = w.state_hash()
    assert c0.count > 3000   # divisions happened
    path = tmp_path / "round_trip.asph"
    w.save_snapshot(path, "unit-test")
    data = path.read_bytes()
    assert data == w.snapshot("unit-test")

    snap = asph.parse(data, ah.CELL_FIELDS)
    h = snap["header"]
    assert (h["magic"], h["version"]) == (asph.MAGIC, asph.VERSION)
    assert (h["seed"], h["tick"], h["sim_time_s"]) == (20260802, c0.tick, c0.sim_time_s)
    assert (h["cell_count"], h["cell_capacity"], h["next_cell_id"]) == (c0.count, 1 << 20, c0.next_id)
    assert (h["taumoeba_count"], h["taumoeba_capacity"]) == (0, 256)   # DEFAULT_TAUMOEBA
    dims = [w.field_dims(f)[0] for f in (ah.TEMPERATURE, ah.CO2, ah.N2)]
    assert [h["n_temp"], h["n_co2"], h["n_n2"]] == [d[0] for d in dims]   # square 2D grids: n
    d = w.desc
    assert (h["chamber_w"], h["chamber_h"], h["chamber_d"]) == (d.chamber_w, d.chamber_h, d.chamber_d)
    assert (h["physics_rate"], h["biology_rate"]) == (1.0, 2.0e7)
    assert h["non_canon_run"] == 1 and h["scenario_id"] == "unit-test"
    assert snap["overrides"] == {"CO2_MASS_PER_DIVISION": 0.5 * CANON_CO2_PER_DIVISION}
    assert h["state_hash"] == asph.fnv1a64(data[asph.HEADER_SIZE:]) == h0
    cells = w.download()
    for k, _ in ah.CELL_FIELDS:
        assert np.array_equal(snap["cells"][k], cells[k]), k
    assert not (cells["flags"] & ah.FLAG_ALIVE).all()   # corpses are serialized too
    for got, want in zip(snap["fields"], world_fields(w)):
        assert np.array_equal(got, want)
    w.close()

    # the file carries the override and the clock; the grids come from the defaults
    with ah.World.load_snapshot(gpu, path, precision=precision) as w2:
        c2 = w2.counters()
        assert (c2.count, c2.tick, c2.next_id, c2.sim_time_s) == (c0.count, c0.tick, c0.next_id, c0.sim_time_s)
        assert w2.desc.co2_mass_per_division == 0.5 * CANON_CO2_PER_DIVISION
        assert w2.desc.biology_rate == 2.0e7 and w2.desc.seed == 20260802
        assert w2.stats()["non_canon_run"] == 1
        assert w2.state_hash() == h0


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_replay_across_restore(gpu, precision):
    """T21.2: the restored world is a continuation point -- 30 ticks past the
    boundary, with divisions, both worlds hold the same full state."""
    a, kw = growth_world(gpu, 2000, precision, absorbing=False)
    a.step(40)
    data = a.snapshot()
    with ah.World.from_snapshot(gpu, data, **kw) as b:   # the motion config is the caller's
        assert b.state_hash() == a.state_hash()
        a.step(30)
        b.step(30)
        ca, cb = a.counters(), b.counters()
        assert ca.count > 2000
        assert (ca.count, ca.next_id, ca.tick) == (cb.count, cb.next_id, cb.tick)
        assert a.state_hash() == b.state_hash()
    a.close()


def test_3d_grids_round_trip_and_replay(gpu):
    """A lit world with 3D temperature and irradiance grids: the header
    carries minus the voxel count, and the continuation is bit-identical."""
    kw = dict(capacity=6000, chamber_w=550e-6, chamber_h=550e-6, chamber_d=100e-6, light_enabled=1,
              light_irradiance=1e4, precision=ah.FP32_HOT,
              grid={0: (64, 64, 16), 1: (16, 16, 4), 2: (-1, 0, 0), 3: (64, 64, 16)})
    a = ah.World(gpu, **kw)
    a.spawn(4000, 5, charge_dist=ah.DIST_UNIFORM, charge_b=0.3, awake=True)
    a.step(6)
    data = a.snapshot("lit")
    h = asph.parse(data, ah.CELL_FIELDS)["header"]
    assert (h["n_temp"], h["n_co2"], h["n_n2"]) == (-64 * 64 * 16, -16 * 16 * 4, 0)
    with ah.World.from_snapshot(gpu, data, **kw) as b:
        assert b.state_hash() == a.state_hash()
        a.step(5)
        b.step(5)
        assert a.state_hash() == b.state_hash()
    a.close()


def test_sticky_non_canon_flag(gpu):
    """A broken canon lock rides the header flag (test_snapshot.cu:71-96): a
    canon world whose snapshot says non-canon stays non-canon."""
    w, kw = growth_world(gpu, 300, ah.FP64_CONTRACT, absorbing=False)
    data = bytearray(w.snapshot())
    assert asph.parse(bytes(data), ah.CELL_FIELDS)["header"]["non_canon_run"] == 0
    assert w.stats()["non_canon_run"] == 0
    data[asph.header_offsets()["non_canon_run"]] = 1   # not hashed: the hash covers the body
    with ah.World.from_snapshot(gpu, bytes(data), **kw) as w2:
        assert w2.stats()["non_canon_run"] == 1
        assert asph.parse(w2.snapshot(), ah.CELL_FIELDS)["header"]["non_canon_run"] == 1
        assert w2.state_hash() == w.state_hash()
    w.close()


def _rehash(data: bytearray) -> bytearray:
    struct.pack_into("<Q", data, asph.header_offsets()["state_hash"], asph.fnv1a64(bytes(data[asph.HEADER_SIZE:])))
    return data


def test_rejects_bad_snapshots(gpu, tmp_path):
    """T21.3 and the loader's other checks, with the reference's status codes
    (src/sim/snapshot.cu:172-224)."""
    grid = {0: (32, 32, 1), 1: (16, 16, 1), 2: (8, 8, 1), 3: (32, 32, 1)}
    w, kw = growth_world(gpu, 500, ah.FP64_CONTRACT, absorbing=False, grid=grid, capacity=1000)
    good = w.snapshot()
    off = asph.header_offsets()

    def status(data, **over):
        with pytest.raises(ah.AstroError) as e:
            ah.World.from_snapshot(gpu, bytes(data), **dict(kw, **over))
        return e.value.status

    bad = bytearray(good)
    struct.pack_into("<I", bad, 0, 0xDEADBEEF)   # clobbered magic
    assert status(bad) == ah.INVALID_ARGUMENT
    bad = bytearray(good)
    struct.pack_into("<I", bad, off["version"], 2)
    assert status(bad) == ah.CONTRACT_VERSION_MISMATCH
    bad = bytearray(good)
    bad[asph.HEADER_SIZE + 1000] ^= 0x10   # a body byte: the hash no longer matches
    assert status(bad) == ah.INVALID_ARGUMENT
    assert status(good[:200]) == ah.FILE_IO_ERROR   # shorter than a header
    assert status(_rehash(bytearray(good[:-4]))) == ah.INVALID_ARGUMENT   # truncated body, consistent hash
    assert status(good, grid={**grid, 0: (64, 64, 1)}) == ah.INVALID_ARGUMENT   # field resolution
    bad = bytearray(good)
    struct.pack_into("<i", bad, off["taumoeba_count"], 1)
    assert status(bad) == ah.INVALID_ARGUMENT   # a predator the body does not hold: truncated
    struct.pack_into("<i", bad, off["taumoeba_count"], 300)
    assert status(bad) == ah.INVALID_ARGUMENT   # more predators than the capacity
    with pytest.raises(ah.AstroError) as e:
        ah.World.load_snapshot(gpu, tmp_path / "missing.asph", **kw)
    assert e.value.status == ah.FILE_IO_ERROR
    with pytest.raises(ah.AstroError) as e:
        w.save_snapshot(tmp_path / "no_such_dir" / "x.asph")
    assert e.value.status == ah.FILE_IO_ERROR
    # a caller buffer below the size
    import ctypes as C
    n = C.c_size_t()
    buf = C.create_string_buffer(16)
    assert gpu.snapshot_to_bytes(w.h, None, C.cast(buf, C.c_void_p), 16, C.byref(n)) == ah.CAPACITY_EXCEEDED
    assert n.value == len(good)
    # the good bytes still load; bytes after the body (hashed) are ignored,
    # as the reference's loader does
    for data in (good, _rehash(bytearray(good + b"\0\0\0\0"))):
        with ah.World.from_snapshot(gpu, bytes(data), **kw) as w2:
            assert w2.state_hash() == w.state_hash()
    w.close()


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_flash_power_override_round_trip(gpu, oracle, precision):
    """PETROVA_FLASH_POWER is a snapshot override like the other two
    (src/sim/snapshot.cu:79-83): saved in the reference's key order, applied
    on load, and the restored world's flash discharges at that power,
    matching the oracle bit for bit."""
    power = 7.5e5
    kw = dict(capacity=400, precision=precision, petrova_flash_power=power, petrova_max_power=2.0e4,
              thermal_enabled=0, grid={0: (64, 64, 1), 1: (32, 32, 1), 2: (16, 16, 1), 3: (64, 64, 1)})
    w = ah.World(gpu, **kw)
    w.spawn(300, 12, charge_dist=ah.DIST_UNIFORM, charge_a=0.2, charge_b=0.9, awake=True)
    data = w.snapshot("flash")
    snap = asph.parse(data, ah.CELL_FIELDS)
    assert list(snap["overrides"].items()) == [("PETROVA_MAX_POWER", 2.0e4), ("PETROVA_FLASH_POWER", power)]
    base = dict(kw)
    base.pop("petrova_flash_power")
    base.pop("petrova_max_power")
    with ah.World.from_snapshot(gpu, data, **base) as r, ah.World(oracle, **kw) as o:
        assert r.desc.petrova_flash_power == power and r.desc.petrova_max_power == 2.0e4
        assert r.state_hash() == w.state_hash()
        o.upload(r.download())
        for x in (r, o):
            x.set_flash(True)
            x.step(2)
        a, b = o.download(), r.download()
        assert np.array_equal(a["energy"], b["energy"])
        ia, ea = o.flash_audit()
        ib, eb = r.flash_audit()
        assert ea == eb and abs(ia - ib) <= 1e-12 * abs(ia)
    w.close()
