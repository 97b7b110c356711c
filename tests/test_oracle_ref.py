"""Pins the CPU oracle to the reference's own code and published vectors.

Three kinds of check:
  * the PCG32 reference vector the reference tests against
    (tests/physics/test_rng.cpp:18-21), hard-coded;
  * bit-for-bit agreement with the reference's host-callable physics headers
    compiled from source into oracle/_ref (rng.cuh, integrator.cuh, thermal.cuh,
    emission.cuh, taxis.cuh, lifecycle.cuh, grid.cuh, hash.cuh, snapshot_v1.h),
    over randomized inputs -- skipped where the reference checkout is absent;
  * the constant table against the reference's generated canon header.
"""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import ah

U64, U32, D, F, I = C.c_uint64, C.c_uint32, C.c_double, C.c_float, C.c_int32
PU64, PD = C.POINTER(C.c_uint64), C.POINTER(C.c_double)


def sig(lib, name, res, args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args
    return fn


@pytest.fixture(scope="module")
def olib(oracle):
    return oracle.lib


def test_pcg32_reference_vector(olib):
    seed = sig(olib, "oracle_pcg32_seed", None, [U64, U64, PU64, PU64])
    nxt = sig(olib, "oracle_pcg32_next", U32, [PU64, U64])
    st, inc = U64(), U64()
    seed(42, 54, C.byref(st), C.byref(inc))
    want = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    got = [nxt(C.byref(st), inc.value) for _ in range(6)]
    assert got == want


def test_constants_match_reference(ref):
    getc = sig(ref, "ref_constant", I, [C.c_char_p, PD])
    import re
    from conftest import ROOT
    text = (ROOT / "include" / "astro_canon.h").read_text()
    defs = dict(re.findall(r"#define ASTRO_([A-Z0-9_]+)\s+\(?([0-9.e+\-u <]+)\)?", text))
    checked = 0
    for name, val in defs.items():
        key = name.replace("FLAG_", "CELL_FLAG_") if name.startswith("FLAG_") else name
        out = D()
        if not getc(key.encode(), C.byref(out)):
            continue
        v = eval(val.replace("u", "")) if "<<" in val else float(val)
        assert float(v) == out.value, (name, v, out.value)
        checked += 1
    assert checked >= 82


def test_rng_streams_bit_exact(olib, ref):
    rng = np.random.default_rng(7)
    pairs = [("splitmix64", [U64]), ("cell_stream_inc", [U64])]
    for name, args in pairs:
        o = sig(olib, "oracle_" + name, U64, args)
        r = sig(ref, "ref_" + name, U64, args)
        for x in rng.integers(0, 2**63, 2000, dtype=np.uint64):
            assert o(int(x)) == r(int(x))
    o = sig(olib, "oracle_cell_rng_init", U64, [U64, U64])
    r = sig(ref, "ref_cell_rng_init", U64, [U64, U64])
    for s, i in zip(rng.integers(0, 2**63, 500, dtype=np.uint64), rng.integers(1, 2**40, 500, dtype=np.uint64)):
        assert o(int(s), int(i)) == r(int(s), int(i))
    o = sig(olib, "oracle_pcg_split", U64, [U64, U64])
    r = sig(ref, "ref_pcg_split", U64, [U64, U64])
    for s, i in zip(rng.integers(0, 2**63, 500, dtype=np.uint64), rng.integers(1, 2**40, 500, dtype=np.uint64)):
        assert o(int(s), int(i)) == r(int(s), int(i))
    # uniform01d and Box-Muller: same bits, same state advance
    ou = sig(olib, "oracle_uniform01d", D, [PU64, U64])
    ru = sig(ref, "ref_uniform01d", D, [PU64, U64])
    og = sig(olib, "oracle_gaussian_pair", None, [PU64, U64, PD, PD])
    rg = sig(ref, "ref_gaussian_pair", None, [PU64, U64, PD, PD])
    inc_fn = sig(olib, "oracle_cell_stream_inc", U64, [U64])
    for s in rng.integers(0, 2**63, 300, dtype=np.uint64):
        inc = inc_fn(int(s) & 0xFFFF)
        a, b = U64(int(s)), U64(int(s))
        assert ou(C.byref(a), inc) == ru(C.byref(b), inc) and a.value == b.value
        z = [D() for _ in range(4)]
        og(C.byref(a), inc, C.byref(z[0]), C.byref(z[1]))
        rg(C.byref(b), inc, C.byref(z[2]), C.byref(z[3]))
        assert (z[0].value, z[1].value, a.value) == (z[2].value, z[3].value, b.value)


def test_motion_physics_bit_exact(olib, ref):
    rng = np.random.default_rng(11)
    for name in ["ou_position_shape", "water_viscosity"]:
        o = sig(olib, "oracle_" + name, D, [D])
        r = sig(ref, "ref_" + name, D, [D])
        xs = rng.uniform(1e-6, 3.0, 500) if name == "ou_position_shape" else rng.uniform(273, 600, 500)
        for x in list(xs) + [1e-2, 1e-2 * (1 - 1e-16), 5e-3, 1e-9]:
            assert o(float(x)) == r(float(x)), (name, x)
    oc = sig(olib, "oracle_ou_coefficients", None, [D, D, D, D, D * 6])
    rc = sig(ref, "ref_ou_coefficients", None, [D, D, D, D, D * 6])
    oa = sig(olib, "oracle_ou_advance_axis", None, [PD, PD, D, D, D * 6, D, I, D, D])
    ra = sig(ref, "ref_ou_advance_axis", None, [PD, PD, D, D, D * 6, D, I, D, D])
    for _ in range(500):
        mass = float(rng.choice([2.1e-14, 1.67e-11, rng.uniform(2.1e-14, 1.7e-11)]))
        gamma = float(rng.uniform(2e-8, 2e-6))
        t = float(rng.uniform(280, 400))
        dt = float(rng.choice([1e-4, 1e-3, 1e-2]))
        a, b = (D * 6)(), (D * 6)()
        oc(mass, gamma, t, dt, a)
        rc(mass, gamma, t, dt, b)
        assert list(a) == list(b)
        p0, v0 = float(rng.uniform(-1e-3, 1e-3)), float(rng.normal(0, 1e-4))
        f = float(rng.normal(0, 1e-10))
        z1, z2 = map(float, rng.normal(size=2))
        noise = int(rng.integers(0, 2))
        po, vo, pr, vr = D(p0), D(v0), D(p0), D(v0)
        oa(C.byref(po), C.byref(vo), f, gamma, a, dt, noise, z1, z2)
        ra(C.byref(pr), C.byref(vr), f, gamma, b, dt, noise, z1, z2)
        assert (po.value, vo.value) == (pr.value, vr.value)


def test_light_thermal_taxis_lifecycle_bit_exact(olib, ref):
    rng = np.random.default_rng(13)
    a = 5.0e-6
    od = sig(olib, "oracle_disc_overlap_fraction", D, [D, D])
    rd = sig(ref, "ref_disc_overlap_fraction", D, [D, D])
    for p in list(rng.uniform(-1e-6, 2.2e-5, 500)) + [0.0, a, 2 * a, 3 * a]:
        assert od(float(p), a) == rd(float(p), a)
    osf = sig(olib, "oracle_shadow_fraction", D, [D * 3, D * 3, D * 3])
    rsf = sig(ref, "ref_shadow_fraction", D, [D * 3, D * 3, D * 3])
    for _ in range(500):
        s = (D * 3)(*rng.normal(0, 1e-5, 3))
        o = (D * 3)(*rng.normal(0, 1e-5, 3))
        ang = rng.uniform(0, 2 * math.pi)
        light = (D * 3)(math.cos(ang), math.sin(ang), 0.0)
        assert osf(s, o, light) == rsf(s, o, light)
    # collinear pair -> exactly 1 (tests/physics/test_emission.cu:63-64)
    assert osf((D * 3)(3 * a, 0, 0), (D * 3)(0, 0, 0), (D * 3)(1, 0, 0)) == 1.0
    ol = sig(olib, "oracle_lumped_exchange_energy", D, [D, D, D, D, D])
    rl = sig(ref, "ref_lumped_exchange_energy", D, [D, D, D, D, D])
    for _ in range(500):
        args = (float(rng.uniform(280, 380)), float(rng.uniform(280, 380)), 3.7573448136933928e-05,
                float(rng.uniform(1e-9, 1e-6)), float(rng.uniform(1e-6, 1e-3)))
        assert ol(*args) == rl(*args)
    oe = sig(olib, "oracle_taxis_ema_update", F, [F, D, D])
    re_ = sig(ref, "ref_taxis_ema_update", F, [F, D, D])
    for _ in range(500):
        args = (float(np.float32(rng.uniform(0, 1e4))), float(rng.uniform(0, 1e4)), float(rng.uniform(1e-5, 1e-2)))
        assert oe(*args) == re_(*args)
    ot = sig(olib, "oracle_taxis_tumble_angle", D, [PU64, U64])
    rt = sig(ref, "ref_taxis_tumble_angle", D, [PU64, U64])
    for s in rng.integers(0, 2**63, 300, dtype=np.uint64):
        x, y = U64(int(s)), U64(int(s))
        assert ot(C.byref(x), 12345 | 1) == rt(C.byref(y), 12345 | 1) and x.value == y.value
    oax = sig(olib, "oracle_division_axis", None, [U64, PD, PD, PD])
    rax = sig(ref, "ref_division_axis", None, [U64, PD, PD, PD])
    for did in rng.integers(1, 2**40, 500, dtype=np.uint64):
        o = [D() for _ in range(3)]
        r = [D() for _ in range(3)]
        oax(int(did), *map(C.byref, o))
        rax(int(did), *map(C.byref, r))
        assert [v.value for v in o] == [v.value for v in r]
    of = sig(olib, "oracle_to_fixed", C.c_int64, [D, D])
    rf = sig(ref, "ref_to_fixed", C.c_int64, [D, D])
    for v in list(rng.normal(0, 1e3, 500)) + [0.5e-9, -0.5e-9, 2.5e-9]:
        assert of(float(v), 1e9) == rf(float(v), 1e9)
    ofn = sig(olib, "oracle_fnv1a64", U64, [C.c_void_p, C.c_size_t, U64])
    rfn = sig(ref, "ref_fnv1a64", U64, [C.c_void_p, C.c_size_t, U64])
    buf = rng.integers(0, 256, 4096, dtype=np.uint8)
    assert ofn(buf.ctypes.data, buf.size, 1469598103934665603) == rfn(buf.ctypes.data, buf.size, 1469598103934665603)


def test_field_sampling_matches_reference(oracle, ref):
    """grid_sample / grid_index_nearest through the oracle's field_sample stage
    against the reference's own functions (src/fields/grid.cuh:109-141)."""
    rs = sig(ref, "ref_grid_sample", F, [C.c_void_p, I, F, F, F])
    rng = np.random.default_rng(17)
    n = 64
    with ah.World(oracle, capacity=600, grid={0: (n, n, 1), 1: (n, n, 1), 3: (n, n, 1)},
                  thermal_enabled=0, emission_enabled=0) as w:
        field = rng.uniform(290, 380, (1, n, n)).astype(np.float32)
        w.set_field(ah.TEMPERATURE, field)
        w.spawn(600, 3)
        cells = w.download()
        # include positions on and beyond the grid edge
        cells["x"][:20] = np.linspace(-0.0021, 0.0021, 20)
        cells["y"][:20] = np.linspace(0.0021, -0.0021, 20)
        w.upload(cells)
        w.stage("forces")
        got = w.download()["t_local"]
        (_, _, _), (dx, _, _) = w.field_dims(ah.TEMPERATURE)
    flat = np.ascontiguousarray(field.ravel())
    want = np.array([rs(flat.ctypes.data, n, float(np.float32(dx)), float(np.float32(x)), float(np.float32(y)))
                     for x, y in zip(cells["x"], cells["y"])], np.float32)
    assert np.array_equal(got, want)


def test_hash_build_matches_reference_keys(oracle, ref):
    """Keys from the reference's hash_key + a stable sort (the CUB contract)
    reproduce the oracle's vals_sorted and bucket ranges exactly."""
    rk = sig(ref, "ref_hash_key", U32, [I, I, I, D, D, D, D])
    with ah.World(oracle, capacity=4000, seed=7) as w:
        w.spawn(4000, 7, placement=ah.PLACE_GAUSSIAN, place_radius=1.5e-4)
        cells = w.download()
        w.hash_build()
        vals, bs, be = w.hash_download()
        nx, ny, nz, cell = w.hash_dims()
    keys = np.array([rk(nx, ny, nz, cell, x, y, z) for x, y, z in zip(cells["x"], cells["y"], cells["z"])],
                    np.uint32)
    order = np.argsort(keys, kind="stable").astype(np.uint32)
    assert np.array_equal(vals, order)
    nb = nx * ny * nz
    sk = keys[order]
    ws, we = np.zeros(nb, np.int32), np.zeros(nb, np.int32)
    uniq, first, counts = np.unique(sk, return_index=True, return_counts=True)
    ws[uniq] = first
    we[uniq] = first + counts
    assert np.array_equal(bs, ws) and np.array_equal(be, we)
