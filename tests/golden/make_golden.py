"""Writes the golden fixtures the oracle tests check against.

    python tests/golden/make_golden.py [--reference /root/reference]

Two files, both data:
  * expected_values.json -- the reference's published physics expectations
    (T1-T26, tests/golden/expected_values.h, values and tolerances), parsed
    from the header's `inline constexpr double NAME = value;` lines;
  * ref_vectors.json -- input/output vectors of the reference's host-callable
    physics (rng.cuh, integrator.cuh, thermal.cuh, emission.cuh, taxis.cuh,
    lifecycle.cuh, fixed_atomic.cuh, hash.cuh, grid.cuh) produced by
    oracle/_ref/libastro_ref.so, which `make -C oracle ref` compiles from the
    reference's own headers. Doubles are stored as float.hex() strings and
    64-bit integers as decimal strings, so every bit survives the round trip.

Run it in a checkout that has the reference; the committed outputs are what
the tests read, so the tests need neither the reference nor oracle/_ref.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import re
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]

U64, U32, D, F, I = C.c_uint64, C.c_uint32, C.c_double, C.c_float, C.c_int32
PU64, PD = C.POINTER(C.c_uint64), C.POINTER(C.c_double)


def parse_expected(header: Path) -> dict:
    pat = re.compile(r"inline\s+constexpr\s+double\s+(\w+)\s*=\s*([-+0-9.eE]+)\s*;")
    out = {}
    for line in header.read_text().splitlines():
        m = pat.search(line)
        if m:
            out[m.group(1)] = float(m.group(2))
    return out


def fn(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def hx(v: float) -> str:
    return float(v).hex()


def ref_vectors(lib) -> dict:
    rng = np.random.default_rng(20260802)
    v: dict[str, list] = {}
    sm = fn(lib, "ref_splitmix64", U64, [U64])
    v["splitmix64"] = [[str(int(x)), str(sm(int(x)))] for x in rng.integers(0, 2**63, 32, dtype=np.uint64)]
    ci = fn(lib, "ref_cell_stream_inc", U64, [U64])
    v["cell_stream_inc"] = [[str(int(x)), str(ci(int(x)))] for x in rng.integers(0, 2**40, 32, dtype=np.uint64)]
    cr = fn(lib, "ref_cell_rng_init", U64, [U64, U64])
    v["cell_rng_init"] = [[str(int(s)), str(int(i)), str(cr(int(s), int(i)))]
                          for s, i in zip(rng.integers(0, 2**63, 32, dtype=np.uint64),
                                          rng.integers(1, 2**40, 32, dtype=np.uint64))]
    ps = fn(lib, "ref_pcg_split", U64, [U64, U64])
    v["pcg_split"] = [[str(int(s)), str(int(i)), str(ps(int(s), int(i)))]
                      for s, i in zip(rng.integers(0, 2**63, 32, dtype=np.uint64),
                                      rng.integers(1, 2**40, 32, dtype=np.uint64))]
    # a stream: 8 uniform01d draws then 4 gaussian pairs from each seed state
    ud = fn(lib, "ref_uniform01d", D, [PU64, U64])
    gp = fn(lib, "ref_gaussian_pair", None, [PU64, U64, PD, PD])
    streams = []
    for s in rng.integers(0, 2**63, 16, dtype=np.uint64):
        inc = ci(int(s) & 0xFFFFF)
        st = U64(int(s))
        draws = [hx(ud(C.byref(st), inc)) for _ in range(8)]
        pairs = []
        for _ in range(4):
            a, b = D(), D()
            gp(C.byref(st), inc, C.byref(a), C.byref(b))
            pairs.append([hx(a.value), hx(b.value)])
        streams.append({"state": str(int(s)), "inc": str(inc), "uniform01d": draws, "gaussian_pairs": pairs,
                        "final_state": str(st.value)})
    v["streams"] = streams
    da = fn(lib, "ref_division_axis", None, [U64, PD, PD, PD])
    axes = []
    for did in rng.integers(1, 2**40, 32, dtype=np.uint64):
        a, b, c = D(), D(), D()
        da(int(did), C.byref(a), C.byref(b), C.byref(c))
        axes.append([str(int(did)), hx(a.value), hx(b.value), hx(c.value)])
    v["division_axis"] = axes
    osh = fn(lib, "ref_ou_position_shape", D, [D])
    xs = list(rng.uniform(1e-6, 3.0, 24)) + [1e-2, 1e-2 * (1 - 1e-16), 5e-3, 1e-9, 50.0]
    v["ou_position_shape"] = [[hx(x), hx(osh(float(x)))] for x in xs]
    wv = fn(lib, "ref_water_viscosity", D, [D])
    v["water_viscosity"] = [[hx(t), hx(wv(float(t)))] for t in list(rng.uniform(273, 600, 24)) + [293.15, 369.565]]
    oc = fn(lib, "ref_ou_coefficients", None, [D, D, D, D, D * 6])
    oa = fn(lib, "ref_ou_advance_axis", None, [PD, PD, D, D, D * 6, D, I, D, D])
    ous = []
    for _ in range(32):
        mass = float(rng.choice([2.1e-14, 1.67e-11, rng.uniform(2.1e-14, 1.7e-11)]))
        gamma = float(rng.uniform(2e-8, 2e-6))
        t = float(rng.uniform(280, 400))
        dt = float(rng.choice([1e-4, 1e-3, 1e-2]))
        c6 = (D * 6)()
        oc(mass, gamma, t, dt, c6)
        p0, v0 = float(rng.uniform(-1e-3, 1e-3)), float(rng.normal(0, 1e-4))
        f = float(rng.normal(0, 1e-10))
        z1, z2 = map(float, rng.normal(size=2))
        noise = int(rng.integers(0, 2))
        p, vv = D(p0), D(v0)
        oa(C.byref(p), C.byref(vv), f, gamma, c6, dt, noise, z1, z2)
        ous.append({"in": [hx(mass), hx(gamma), hx(t), hx(dt)], "coeff": [hx(x) for x in c6],
                    "advance_in": [hx(p0), hx(v0), hx(f), hx(z1), hx(z2), noise], "advance_out": [hx(p.value), hx(vv.value)]})
    v["ou"] = ous
    a = 5.0e-6
    dof = fn(lib, "ref_disc_overlap_fraction", D, [D, D])
    ps_ = list(rng.uniform(0, 2.2 * a, 24)) + [0.0, 2 * a, 2 * a * (1 - 1e-15), 1e-12]
    v["disc_overlap_fraction"] = [[hx(p), hx(a), hx(dof(float(p), a))] for p in ps_]
    sf = fn(lib, "ref_shadow_fraction", D, [D * 3, D * 3, D * 3])
    shadows = []
    for _ in range(24):
        s3 = (D * 3)(*rng.uniform(-1e-4, 1e-4, 3))
        o3 = (D * 3)(*(np.array(s3) + rng.normal(0, 6e-6, 3)))
        l3 = rng.normal(size=3)
        l3 = (D * 3)(*(l3 / np.linalg.norm(l3)))
        shadows.append([[hx(x) for x in s3], [hx(x) for x in o3], [hx(x) for x in l3], hx(sf(s3, o3, l3))])
    v["shadow_fraction"] = shadows
    le = fn(lib, "ref_lumped_exchange_energy", D, [D, D, D, D, D])
    lum = []
    for _ in range(24):
        args = [float(rng.uniform(280, 380)), float(rng.uniform(280, 380)), float(rng.uniform(1e-6, 1e-4)),
                float(rng.choice([1e-9, 1e-6, 4e-3])), float(rng.choice([1e-5, 1e-4, 1e-3]))]
        lum.append([*(hx(x) for x in args), hx(le(*args))])
    v["lumped_exchange_energy"] = lum
    te = fn(lib, "ref_taxis_ema_update", F, [F, D, D])
    ema = []
    for _ in range(24):
        e0, s, dt = float(np.float32(rng.uniform(0, 1e4))), float(rng.uniform(0, 1e4)), float(rng.choice([1e-4, 1e-3, 1e-2]))
        ema.append([hx(e0), hx(s), hx(dt), hx(te(e0, s, dt))])
    v["taxis_ema_update"] = ema
    tt = fn(lib, "ref_taxis_tumble_angle", D, [PU64, U64])
    tum = []
    for s in rng.integers(0, 2**63, 16, dtype=np.uint64):
        inc = ci(int(s) & 0xFFFF)
        st = U64(int(s))
        ang = [hx(tt(C.byref(st), inc)) for _ in range(4)]
        tum.append({"state": str(int(s)), "inc": str(inc), "angles": ang, "final_state": str(st.value)})
    v["taxis_tumble_angle"] = tum
    tf = fn(lib, "ref_to_fixed", C.c_int64, [D, D])
    fx = []
    for _ in range(24):
        x, sc = float(rng.normal(0, 1e3)), float(rng.choice([1e6, 1e9, 1e15, 1e18]))
        fx.append([hx(x * 1e-6), hx(sc), str(tf(x * 1e-6, sc))])
    v["to_fixed"] = fx
    hk = fn(lib, "ref_hash_key", U32, [I, I, I, D, D, D, D])
    keys = []
    for _ in range(32):
        nx, ny, nz = map(int, rng.integers(4, 400, 3))
        cell = float(rng.choice([1e-5, 2.2e-5]))
        p3 = [float(rng.uniform(-0.6, 0.6) * n * cell) for n in (nx, ny, nz)]
        keys.append([nx, ny, nz, hx(cell), *(hx(x) for x in p3), int(hk(nx, ny, nz, cell, *p3))])
    v["hash_key"] = keys
    # the predator's helpers (src/sim/predation.cuh:24-80)
    td = fn(lib, "ref_tau_drag", D, [D])
    v["tau_drag"] = [[hx(t), hx(td(float(t)))] for t in list(rng.uniform(273, 400, 12)) + [293.15]]
    to = fn(lib, "ref_tau_overlaps_cell", I, [D * 3, D * 3])
    reach = 2.0000000000000002e-05 + 5.0000000000000004e-06
    ov = []
    for k in range(24):
        a3 = rng.uniform(-1e-3, 1e-3, 3)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        dist = reach * (1.0 + (k - 12) * 1e-15) if k < 24 else reach
        b3 = a3 + u * (dist if k % 3 else float(rng.uniform(0, 2 * reach)))
        ov.append([[hx(x) for x in a3], [hx(x) for x in b3], int(to((D * 3)(*a3), (D * 3)(*b3)))])
    v["tau_overlaps_cell"] = ov
    tt_ = fn(lib, "ref_tau_should_tumble", I, [F, C.POINTER(F), F, D])
    tum2 = []
    for _ in range(24):
        sig_ = float(np.float32(rng.integers(0, 40)))
        e0 = float(np.float32(rng.uniform(0, 40)))
        run = float(np.float32(rng.choice([0.0, 0.1, 0.39, 0.4, 0.41])))
        dt = float(rng.choice([1e-4, 1e-3, 1e-2]))
        e = F(e0)
        r = tt_(sig_, C.byref(e), run, dt)
        tum2.append([hx(sig_), hx(e0), hx(run), hx(dt), int(r), hx(e.value)])
    v["tau_should_tumble"] = tum2
    th = fn(lib, "ref_tau_n2_hazard", D, [D, D])
    v["tau_n2_hazard"] = [[hx(n), hx(t), hx(th(float(n), float(t)))]
                          for n, t in zip(rng.uniform(0, 0.03, 16), rng.uniform(0, 1, 16))]
    tp = fn(lib, "ref_tau_n2_death_prob", D, [D, D])
    v["tau_n2_death_prob"] = [[hx(h), hx(d), hx(tp(float(h), float(d)))]
                              for h, d in zip(list(rng.uniform(0, 0.02, 14)) + [0.0, 1e-9],
                                              list(rng.uniform(0, 500, 14)) + [1.0, 0.5])]
    tdt = fn(lib, "ref_tau_daughter_tolerance", F, [F, D])
    v["tau_daughter_tolerance"] = [[hx(float(np.float32(p_))), hx(m), hx(tdt(float(np.float32(p_)), float(m)))]
                                   for p_, m in zip(list(rng.uniform(0, 1, 14)) + [0.0, 1.0],
                                                    list(rng.normal(0, 3, 14)) + [-60.0, 60.0])]
    return v


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    a = ap.parse_args()
    ref = Path(a.reference)
    exp = parse_expected(ref / "tests" / "golden" / "expected_values.h")
    (HERE / "expected_values.json").write_text(json.dumps(
        {"source": "tests/golden/expected_values.h (generated there by scripts/derive.py)", "values": exp},
        indent=1, sort_keys=True) + "\n")
    subprocess.run(["make", "-C", str(ROOT / "oracle"), f"REF={ref}", "ref"], check=True, stdout=subprocess.DEVNULL)
    lib = C.CDLL(str(ROOT / "oracle" / "_ref" / "libastro_ref.so"))
    vec = ref_vectors(lib)
    (HERE / "ref_vectors.json").write_text(json.dumps(
        {"source": "oracle/_ref/libastro_ref.so built from the reference's headers", "vectors": vec},
        indent=0, sort_keys=True) + "\n")
    print(f"{len(exp)} expected values, {sum(len(x) for x in vec.values())} reference vectors")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
