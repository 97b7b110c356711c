"""The allocation-free hash build (astro_hash_build_noalloc) against the
world's own hash build (astro_stage_hash_build + astro_hash_download) on the
same cells: vals_sorted and every bucket range equal, on two calls in a row,
the second smaller; repeated calls take no device memory; sizes beyond the
stated maxima are refused before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from conftest import ah

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """Device buffers for the caller's side of the call (the test may
    allocate; the entry point under test may not)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.ptrs = []

    def put(self, a: np.ndarray) -> C.c_void_p:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 4))) == 0
        self.ptrs.append(p)
        if a.nbytes:
            assert self.lib.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0
        return p

    def empty(self, n: int, dtype) -> C.c_void_p:
        return self.put(np.zeros(n, dtype))

    def get(self, p: C.c_void_p, n: int, dtype) -> np.ndarray:
        a = np.zeros(n, dtype)
        assert self.lib.hipDeviceSynchronize() == 0
        if n:
            assert self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, C.c_size_t(a.nbytes), D2H) == 0
        return a

    def free_memory(self) -> int:
        assert self.lib.hipDeviceSynchronize() == 0
        f, t = C.c_size_t(), C.c_size_t()
        assert self.lib.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    def release(self):
        for p in self.ptrs:
            self.lib.hipFree(p)
        self.ptrs = []


def cells_for(n: int, chamber, seed: int) -> dict:
    rng = np.random.default_rng(seed)
    w, h, d = chamber
    m = n // 3
    x = np.concatenate([rng.uniform(-0.5, 0.5, n - m) * w, rng.normal(0.0, 2e-5, m)])
    y = np.concatenate([rng.uniform(-0.5, 0.5, n - m) * h, rng.normal(1e-4, 2e-5, m)])
    z = rng.uniform(-0.5, 0.5, n) * d
    flags = np.full(n, ah.FLAG_OCCUPIED | ah.FLAG_ALIVE, np.uint32)
    flags[rng.random(n) < 0.03] = 0
    flags[rng.random(n) < 0.05] &= ~np.uint32(ah.FLAG_ALIVE)
    return {"id": np.arange(1, n + 1, dtype=np.uint64), "flags": flags, "x": x, "y": y, "z": z}


def world_hash(gpu, cells: dict, chamber, hash_cell: float):
    n = len(cells["id"])
    with ah.World(gpu, capacity=n, chamber_w=chamber[0], chamber_h=chamber[1], chamber_d=chamber[2],
                  hash_cell_size=hash_cell,
                  grid={ah.TEMPERATURE: (16, 16, 1), ah.CO2: (8, 8, 1), ah.N2: (-1, 0, 0),
                        ah.IRRADIANCE: (8, 8, 1)}) as w:
        w.upload(cells)
        w.hash_build()
        vals, bs, be = w.hash_download()
        dims = w.hash_dims()
    return vals, bs, be, dims


def noalloc_hash(gpu, hip: Hip, cells: dict, dims):
    nx, ny, nz, cell = dims
    n, nb = len(cells["id"]), nx * ny * nz
    args = [hip.put(cells[k].astype(np.float64)) for k in ("x", "y", "z")]
    flags = hip.put(cells["flags"])
    vals, bs, be = hip.empty(n, np.uint32), hip.empty(nb, np.int32), hip.empty(nb, np.int32)
    st = gpu.hash_build_noalloc(*args, flags, n, nx, ny, nz, cell, vals, bs, be, None)
    assert st == ah.OK, ah.STATUS[st]
    return hip.get(vals, n, np.uint32), hip.get(bs, nb, np.int32), hip.get(be, nb, np.int32)


def test_noalloc_hash_equals_world_hash(gpu):
    hip = Hip()
    try:
        # 1M cells at 10 um in the 4 x 4 x 0.2 mm chamber (400 x 400 x 20
        # buckets), then a small chamber at 22 um: the second call reuses the
        # scratch the first left behind.
        cases = [(1_000_000, (4e-3, 4e-3, 2e-4), 1e-5, 1), (2_000, (5.5e-4, 5.5e-4, 1e-4), 0.0, 2),
                 (0, (5.5e-4, 5.5e-4, 1e-4), 0.0, 3)]
        for k, (n, chamber, hash_cell, seed) in enumerate(cases):
            cells = cells_for(n, chamber, seed)
            want = world_hash(gpu, cells, chamber, hash_cell)
            got = noalloc_hash(gpu, hip, cells, want[3])
            print(f"call {k}: n={n} buckets={np.prod(want[3][:3])}", flush=True)
            for name, a, b in zip(("vals_sorted", "bucket_start", "bucket_end"), want[:3], got):
                assert np.array_equal(a, b), (k, name, np.count_nonzero(a != b))
            assert n == 0 or np.count_nonzero(got[2] > got[1]) > 100
    finally:
        hip.release()


def test_noalloc_hash_takes_no_memory(gpu):
    hip = Hip()
    try:
        chamber = (1e-3, 1e-3, 2e-4)
        cells = cells_for(50_000, chamber, 4)
        _, _, _, dims = world_hash(gpu, cells, chamber, 1e-5)
        nx, ny, nz, cell = dims
        n, nb = len(cells["id"]), nx * ny * nz
        xyz = [hip.put(cells[k]) for k in ("x", "y", "z")]
        flags = hip.put(cells["flags"])
        out = [hip.empty(n, np.uint32), hip.empty(nb, np.int32), hip.empty(nb, np.int32)]
        assert gpu.hash_build_noalloc(*xyz, flags, n, nx, ny, nz, cell, *out, None) == ah.OK
        before = hip.free_memory()
        for _ in range(50):
            assert gpu.hash_build_noalloc(*xyz, flags, n, nx, ny, nz, cell, *out, None) == ah.OK
        after = hip.free_memory()
        assert before - after < (1 << 20), before - after
        # limits: refused before any launch
        assert gpu.hash_build_noalloc(*xyz, flags, ah_max_cells() + 1, nx, ny, nz, cell, *out, None) \
            == ah.CAPACITY_EXCEEDED
        assert gpu.hash_build_noalloc(*xyz, flags, n, 2048, 2048, 2, cell, *out, None) == ah.CAPACITY_EXCEEDED
        assert gpu.hash_build_noalloc(*xyz, flags, -1, nx, ny, nz, cell, *out, None) == ah.INVALID_ARGUMENT
        assert gpu.hash_build_noalloc(*xyz, flags, n, nx, ny, nz, 0.0, *out, None) == ah.INVALID_ARGUMENT
    finally:
        hip.release()


def ah_max_cells() -> int:
    return 1 << 20
