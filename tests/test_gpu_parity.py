"""GPU parity: the HIP library against the CPU oracle on the same inputs.

Bar: integer state (ids, flags, death causes, RNG states, slot order, hash
order and bucket ranges, fixed-point deposits) is bit-exact. FP64-layout
floating state agrees to a few ulp: the only sources of difference are the
device's and glibc's transcendental functions (exp, expm1, log, sin, cos,
acos), which are each within an ulp or two of exact. FP32-layout positions are
FP64 arithmetic rounded to float on store, checked against the FP64 oracle
within the stated absolute tolerances.
"""
import numpy as np
import pytest

from conftest import ah, compare_cells, oracle_cells

pytestmark = pytest.mark.gpu

C1 = dict(chamber_w=550e-6, chamber_h=550e-6, chamber_d=100e-6, physics_rate=0.1,
          grid={0: (64, 64, 64), 1: (64, 64, 64), 2: (-1, 0, 0), 3: (64, 64, 64)})
REF2D = dict(grid={0: (128, 128, 1), 1: (64, 64, 1), 2: (32, 32, 1), 3: (128, 128, 1)})


def pair(oracle, gpu, cells=None, **kw):
    o = ah.World(oracle, **kw)
    g = ah.World(gpu, **kw)
    if cells is not None:
        o.upload(cells)
        g.upload(cells)
    return o, g


def rel_close(a, b, rtol, atol=0.0, name=""):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(a).max() if a.size else 0.0, 1e-300)
    d = np.abs(a - b).max() if a.size else 0.0
    assert d <= rtol * scale + atol, f"{name}: max diff {d:.3e} vs scale {scale:.3e}"


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_spawn_parity(oracle, gpu, precision):
    kw = dict(C1, capacity=2000, precision=precision)
    o, g = pair(oracle, gpu, **kw)
    spawn = dict(charge_dist=ah.DIST_UNIFORM, charge_a=0.0, charge_b=0.2)
    o.spawn(2000, 1, **spawn)
    g.spawn(2000, 1, **spawn)
    a, b = o.download(), g.download()
    compare_cells(a, b)
    for name in ("x", "y", "z", "energy"):
        want = a[name] if precision == ah.FP64_CONTRACT else a[name].astype(np.float32).astype(np.float64)
        if name == "energy":
            want = a[name]
        assert np.array_equal(want, b[name]), name
    for name in ("dir_x", "dir_y", "dir_z"):
        rel_close(a[name], b[name], 1e-6, name=name)
    assert o.counters().next_id == g.counters().next_id == 2001


@pytest.mark.parametrize("case", ["c1", "cluster", "c3"])
def test_hash_parity(oracle, gpu, case):
    if case == "c1":
        kw, n, spawn = dict(C1, capacity=2000), 2000, {}
    elif case == "cluster":
        kw, n, spawn = dict(capacity=4000, seed=7), 4000, dict(placement=ah.PLACE_GAUSSIAN, place_radius=1.5e-4)
    else:
        kw = dict(capacity=1_000_000, chamber_w=4e-3, chamber_h=4e-3, chamber_d=2e-4, hash_cell_size=1e-5,
                  grid={0: (64, 64, 1), 1: (32, 32, 1), 2: (-1, 0, 0), 3: (64, 64, 1)})
        n, spawn = 1_000_000, {}
    cells = oracle_cells(oracle, n, 1, spawn=spawn, **kw)
    # a few unoccupied slots exercise the sentinel bucket
    cells["flags"][::997] = 0
    for precision in (ah.FP64_CONTRACT, ah.FP32_HOT):
        kwp = dict(kw, precision=precision)
        c = dict(cells)
        if precision == ah.FP32_HOT:
            for k in ("x", "y", "z"):
                c[k] = c[k].astype(np.float32).astype(np.float64)
        o, g = pair(oracle, gpu, c, **kwp)
        o.hash_build()
        g.hash_build()
        vo, so, eo = o.hash_download()
        vg, sg, eg = g.hash_download()
        assert np.array_equal(vo, vg), f"vals_sorted differs at {np.nonzero(vo != vg)[0][:10]}"
        assert np.array_equal(so, sg) and np.array_equal(eo, eg)
        # rebuilding is idempotent and order-stable (tests/physics/test_hash.cu:138-158)
        g.hash_build()
        assert np.array_equal(g.hash_download()[0], vg)
        o.close()
        g.close()


def staging_demand(w) -> int:
    """Largest number of records one wave of the contact walk stages for one
    z offset: the 64 consecutive sorted records of a wave have bucket keys in
    [kmin, kmax], and row (dz, dy) of every lane lies in buckets
    [kmin + o - 1, kmax + o + 1], o = dz*nx*ny + dy*nx (clamped)."""
    nx, ny, nz, _ = w.hash_dims()
    nb = nx * ny * nz
    _, bs, be = w.hash_download()
    counts = (be - bs).astype(np.int64)
    bstart = np.concatenate([[0], np.cumsum(counts)])
    keys = np.repeat(np.arange(nb), counts)
    worst = 0
    for w0 in range(0, keys.size, 64):
        kmin, kmax = int(keys[w0:w0 + 64].min()), int(keys[w0:w0 + 64].max())
        for dz in (-1, 0, 1):
            tot = 0
            for dy in (-1, 0, 1):
                o = dz * nx * ny + dy * nx
                lo = bstart[min(max(kmin + o - 1, 0), nb)]
                hi = bstart[min(max(kmax + o + 2, 0), nb)]
                tot += max(int(hi - lo), 0)
            worst = max(worst, tot)
    return worst


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
@pytest.mark.parametrize("dims", ["2d", "3d"])
def test_contact_forces_dense_cluster(oracle, gpu, precision, dims):
    """Contact forces of a cluster so dense that a wave's rows of one z
    offset hold over twice the records its LDS slice takes (256 float or 192
    double records), so the walk stages them in three or more chunks: the
    forces are the oracle's bit for bit (FP64 layout) or to float rounding of
    the body force (FP32 layout)."""
    if dims == "2d":
        kw = dict(REF2D, capacity=6000, hash_cell_size=1e-5, precision=precision)
        n, radius = 6000, 28e-6
    else:
        kw = dict(C1, capacity=8000, hash_cell_size=1e-5, precision=precision)
        n, radius = 8000, 15e-6
    cells = oracle_cells(oracle, n, 9, spawn=dict(placement=ah.PLACE_GAUSSIAN, place_radius=radius), **kw)
    cells["flags"][::501] = 0
    if precision == ah.FP32_HOT:
        for k in ("x", "y", "z"):
            cells[k] = cells[k].astype(np.float32).astype(np.float64)
    o, g = pair(oracle, gpu, cells, **kw)
    with o, g:
        o.hash_build()
        assert staging_demand(o) > 2 * 256   # three or more chunks in either layout
        g.hash_build()
        o.stage("forces")
        g.stage("forces")
        fo, fg = o.forces(), g.forces()
        for a, b, name in zip(fo, fg, "xyz"):
            assert np.abs(a).max() > 0.0
            if precision == ah.FP64_CONTRACT:
                assert np.array_equal(a, b), name
        if precision == ah.FP32_HOT:
            forces_close_f32(fo, fg, cells, kw.get("physics_rate", 1.0))


# The contact diameter 2a (ASTRO_CELL_DIAMETER, include/astro_canon.h) and
# the canon contact stiffness (ASTRO_CONTACT_STIFFNESS).
TWO_A = 1.0000000000000001e-05
K_CONTACT = 1.1802e-05


def contact_counts(cells: dict) -> np.ndarray:
    """Occupied neighbours of each cell within 2a (brute force), counting
    those a few float roundings beyond it too, which the float test may take
    as touching."""
    p = np.stack([cells["x"], cells["y"], cells["z"]], axis=1)
    occ = (cells["flags"] & ah.FLAG_OCCUPIED) != 0
    out = np.zeros(len(p), np.int64)
    q = p[occ]
    for s0 in range(0, len(p), 512):
        d2 = ((p[s0:s0 + 512, None, :] - q[None, :, :]) ** 2).sum(axis=2)
        out[s0:s0 + 512] = ((d2 < (TWO_A * (1.0 + 2.0 ** -19)) ** 2) & (d2 > 0.0)).sum(axis=1)
    return out


def f32_force_bound(f_oracle: np.ndarray, contacts: np.ndarray, physics_rate: float) -> np.ndarray:
    """Per-cell bound on |float force - double force| in the FP32 layout,
    which sums contact forces in float. One term k (2a - r) d/r, with r from
    float differences, squares and a square root, is within about
    k * 2a * 2^-21 of the double one, and each float add rounds by at most
    2^-24 of the running sum, itself at most n * k * 2a for n contacts (each
    term is at most k * 2a); the body force is rounded to float once. For a
    cell in contact with a few others at dt = 1e-4 s (k = 1.2e-4 N/m) this is
    ~1e-17 N, moving it by ~1e-14 m per tick, far inside FP32_POS_TOL."""
    k = K_CONTACT / physics_rate
    n = contacts.astype(np.float64)
    return 2.0 ** -21 * n * k * TWO_A + 2.0 ** -23 * n * n * k * TWO_A + 2.0 ** -23 * np.abs(f_oracle)


def forces_close_f32(fo, fg, cells, physics_rate):
    bound = f32_force_bound(np.max(np.abs(np.stack(fo)), axis=0), contact_counts(cells), physics_rate)
    for a, b, name in zip(fo, fg, "xyz"):
        bad = np.nonzero(np.abs(a - b) > bound)[0]
        assert bad.size == 0, (name, bad[:5], np.abs(a - b)[bad[:5]], bound[bad[:5]])


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_contact_screen_threshold_pairs(oracle, gpu, precision):
    """Isolated pairs of cells whose separations lie within 2e-6 (relative)
    of the contact diameter, on either side of it and in random directions:
    the walk's float pre-screen must keep every pair the double test counts
    as touching, so the forces are the oracle's bit for bit (FP64 layout) or
    within f32_force_bound (FP32 layout)."""
    kw = dict(C1, capacity=1000, hash_cell_size=1e-5, precision=precision)
    xs = -240e-6 + 40e-6 * np.arange(13)
    centres = np.array([(x, y, z) for z in (-25e-6, 25e-6) for y in xs for x in xs])
    npair = centres.shape[0]
    cells = oracle_cells(oracle, 2 * npair, 3, **kw)
    rng = np.random.default_rng(11)
    rel = np.concatenate([np.linspace(-2e-6, 2e-6, npair // 2), rng.uniform(-2e-6, 2e-6, npair - npair // 2)])
    dirs = rng.normal(size=(npair, 3))
    dirs[:6] = np.eye(3)[[0, 1, 2, 0, 1, 2]]   # a few along the axes
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    half = 0.5 * TWO_A * (1.0 + rel)[:, None] * dirs
    pos = np.empty((2 * npair, 3))
    pos[0::2] = centres + half
    pos[1::2] = centres - half
    if precision == ah.FP32_HOT:
        pos = pos.astype(np.float32).astype(np.float64)
    for j, k in enumerate(("x", "y", "z")):
        cells[k] = pos[:, j].copy()
    r = np.linalg.norm(pos[0::2] - pos[1::2], axis=1)
    near = np.abs(r / TWO_A - 1.0) < 3e-6
    assert near.all()
    assert (r < TWO_A).sum() > 100 and (r >= TWO_A).sum() > 100   # both sides of the threshold
    o, g = pair(oracle, gpu, cells, **kw)
    with o, g:
        o.hash_build()
        g.hash_build()
        o.stage("forces")
        g.stage("forces")
        fo, fg = o.forces(), g.forces()
        if precision == ah.FP64_CONTRACT:
            for a, b, name in zip(fo, fg, "xyz"):
                assert np.array_equal(a, b), name
        else:
            # a pair within a few float roundings of 2a may count on one side
            # only: one contact each way
            forces_close_f32(fo, fg, cells, kw.get("physics_rate", 1.0))


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
@pytest.mark.parametrize("light", [(1.0, 0.0), (0.6, 0.8)])
def test_shadow_screen_threshold_pairs(oracle, gpu, precision, light):
    """Isolated pairs, an awake cell and an upstream neighbour placed at the
    edges of the shadow walk's float screen: half at a perpendicular offset
    within 1e-4..1e-3 (relative) of 2a, inside and outside; half 5 um off the
    light line with an along-light offset within +-1e-7..1e-5 of zero. On the
    oracle the near-field product moves the shaded cell's irradiance by at
    least 12 float ulps in every shading case and by none otherwise, so
    agreement to two ulps per cell means the screen kept every shading
    neighbour and no other."""
    kw = dict(C1, capacity=1000, hash_cell_size=1e-5, thermal_enabled=0, light_enabled=1, light_irradiance=1e4,
              light_dir_x=light[0], light_dir_y=light[1], precision=precision)
    xs = -240e-6 + 40e-6 * np.arange(13)
    centres = np.array([(x, y, z) for z in (-25e-6, 25e-6) for y in xs for x in xs])
    npair = centres.shape[0]
    cells = oracle_cells(oracle, 2 * npair, 3, spawn=dict(awake=True, charge_dist=ah.DIST_UNIFORM, charge_b=0.9),
                         **kw)
    ld = np.array([light[0], light[1], 0.0])
    ld /= np.linalg.norm(ld)
    across = np.array([-ld[1], ld[0], 0.0])
    h = npair // 2
    q = (npair - h) // 2
    perp = np.concatenate([TWO_A * (1 - np.logspace(-4, -3, h // 2)), TWO_A * (1 + np.logspace(-4, -3, h - h // 2)),
                           np.full(npair - h, 5e-6)])
    along = np.concatenate([np.full(h, 6e-6), 5e-6 * np.logspace(-7, -5, q), -5e-6 * np.logspace(-7, -5, npair - h - q)])
    pos = np.empty((2 * npair, 3))
    pos[0::2] = centres + 0.5 * along[:, None] * ld                          # the shaded cell
    pos[1::2] = centres - 0.5 * along[:, None] * ld + perp[:, None] * across  # its upstream neighbour
    if precision == ah.FP32_HOT:
        pos = pos.astype(np.float32).astype(np.float64)
    for j, k in enumerate(("x", "y", "z")):
        cells[k] = pos[:, j].copy()
    o, g = pair(oracle, gpu, cells, **kw)
    with o, g:
        o.step(1)
        g.step(1)
        a, b = o.download()["irradiance"], g.download()["irradiance"]
    ulp = np.spacing(np.abs(a).astype(np.float32)).astype(np.float64)
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    bad = np.nonzero(d > 2 * ulp)[0]
    assert bad.size == 0, f"irradiance differs at cells {bad[:10]}: {a[bad[:5]]} vs {b[bad[:5]]}"


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_hash_bucket_edges(oracle, gpu, precision):
    """Cells on bucket edges and up to two ulps (of the layout's coordinate
    type) to either side. The hash keeps its origin and inverse cell size as
    floats (the reference's geometry), so an edge sits at
    float(origin) + k / float(1 / cell), not at origin + k * cell; positions
    within two ulps of that straddle it. The bucket of each cell, and so the
    sorted slots and bucket ranges, are the oracle's bit for bit."""
    kw = dict(C1, capacity=2000, hash_cell_size=1e-5, precision=precision)
    n = 1500
    cells = oracle_cells(oracle, n, 5, **kw)
    with ah.World(oracle, **kw) as w:
        nx, ny, nz, cell = w.hash_dims()
    dims = np.array([nx, ny, nz])
    inv = np.float64(np.float32(1.0 / cell))
    rng = np.random.default_rng(17)
    ftype = np.float32 if precision == ah.FP32_HOT else np.float64
    pos = np.empty((n, 3))
    straddle = 0
    for a in range(3):
        origin = np.float64(np.float32(-0.5 * dims[a] * cell))
        k = rng.integers(1, dims[a], n)
        v = (origin + k / inv).astype(ftype)
        step = rng.integers(-2, 3, n)
        for _ in range(2):
            v = np.where(step > 0, np.nextafter(v, ftype(np.inf)), np.where(step < 0, np.nextafter(v, ftype(-np.inf)), v))
            step = step - np.sign(step)
        pos[:, a] = v.astype(np.float64)
        straddle += np.count_nonzero(np.trunc((pos[:, a] - origin) * inv) == k - 1)
    assert straddle > n // 4   # a good share of the cells fall just below an edge
    for j, k in enumerate(("x", "y", "z")):
        cells[k] = pos[:, j].copy()
    o, g = pair(oracle, gpu, cells, **kw)
    with o, g:
        o.hash_build()
        g.hash_build()
        vo, so, eo = o.hash_download()
        vg, sg, eg = g.hash_download()
        assert np.array_equal(vo, vg), f"vals_sorted differs at {np.nonzero(vo != vg)[0][:10]}"
        assert np.array_equal(so, sg) and np.array_equal(eo, eg)


@pytest.mark.parametrize("dims,bc_field", [((128, 128, 1), ah.TEMPERATURE), ((96, 80, 24), ah.TEMPERATURE),
                                           ((64, 64, 16), ah.CO2)])
def test_ftcs_parity(oracle, gpu, dims, bc_field):
    kw = dict(capacity=16, grid={0: dims, 1: dims if bc_field == ah.CO2 else (32, 32, 1), 2: (-1, 0, 0),
                                 3: (32, 32, 1)})
    o, g = pair(oracle, gpu, **kw)
    rng = np.random.default_rng(5)
    nx, ny, nz = dims
    v = rng.uniform(280.0, 400.0, (nz, ny, nx)).astype(np.float32)
    o.set_field(bc_field, v)
    g.set_field(bc_field, v)
    for _ in range(7):
        o.stage("diffuse", bc_field, 1e-3)
        g.stage("diffuse", bc_field, 1e-3)
    assert np.array_equal(o.field(bc_field), g.field(bc_field))


# FP32 storage: the integrator's drag, coefficients and Gaussians run in
# float (about 1e-7 of a displacement) and each tick rounds positions to
# float, so a coordinate may land one ulp (2.3e-10 m at 2 mm) from the FP64
# oracle's rounding of it on some ticks; two ulps bound it over these runs.
FP32_POS_TOL = 5.0e-10
# A one-ulp position difference inside a contact changes the contact force by
# k * 1e-10 m and the drift velocity by that over gamma: ~1.3e-7 m/s per ulp
# at the C1 stiffness (1.18e-4 N/m, gamma ~ 9e-8 kg/s).
FP32_VEL_TOL = 5.0e-7


def test_tick_parity_fp32_dormant(oracle, gpu):
    """C1 shape in FP32 storage, 2000 dormant cells, 10 ticks: integer state
    (flags, RNG streams) bit-exact, positions within FP32_POS_TOL."""
    kw = dict(C1, capacity=2000, ambient_temp=369.0, precision=ah.FP32_HOT)
    cells = oracle_cells(oracle, 2000, 2, spawn=dict(charge_dist=ah.DIST_UNIFORM, charge_b=0.2), **kw)
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        cells[k] = cells[k].astype(np.float32).astype(np.float64)
    o, g = pair(oracle, gpu, cells, **kw)
    o.step(10)
    g.step(10)
    a, b = o.download(), g.download()
    compare_cells(a, b)
    for name in ("x", "y", "z"):
        rel_close(a[name], b[name], 0.0, atol=FP32_POS_TOL, name=name)
    for name in ("vx", "vy", "vz"):
        rel_close(a[name], b[name], 0.0, atol=FP32_VEL_TOL, name=name)


def test_tick_parity_fp64_dormant(oracle, gpu):
    """C1 shape, 2000 dormant cells, T 64^3 at 369.0 K (below ignition)."""
    kw = dict(C1, capacity=2000, ambient_temp=369.0)
    cells = oracle_cells(oracle, 2000, 1, spawn=dict(charge_dist=ah.DIST_UNIFORM, charge_b=0.2), **kw)
    o, g = pair(oracle, gpu, cells, **kw)
    for ticks in (1, 9):
        o.step(ticks)
        g.step(ticks)
        a, b = o.download(), g.download()
        compare_cells(a, b)
        for name in ("x", "y", "z", "vx", "vy", "vz", "energy"):
            rel_close(a[name], b[name], 1e-12, name=name)
        rel_close(a["t_local"], b["t_local"], 1e-6, name="t_local")
        rel_close(o.field(ah.TEMPERATURE), g.field(ah.TEMPERATURE), 1e-6, name="T")
    so, sg = o.stats(), g.stats()
    assert so["n_live"] == sg["n_live"] == 2000
    assert abs(so["total_energy_j"] - sg["total_energy_j"]) <= 1e-6 * so["total_energy_j"]


@pytest.mark.parametrize("light", [(1.0, 0.2), (-1.0, 0.3), (0.2, 1.0), (0.3, -0.9)])
def test_tick_parity_awake_lit(oracle, gpu, light):
    """Reference 2D geometry with awake cells, light and taxis on: thermostat
    deposits, occlusion sweep, exact near-field shadow, feeding, tumbles.
    The four light directions sweep along +x, -x, +y and -y; after the first
    tick (identical positions) the swept irradiance field is bit-exact."""
    kw = dict(REF2D, capacity=3000, chamber_w=1e-3, chamber_h=1e-3, ambient_temp=300.0, light_enabled=1,
              light_irradiance=1.0e4, light_dir_x=light[0], light_dir_y=light[1], co2_init=1e-3, biology_rate=1e6)
    cells = oracle_cells(oracle, 3000, 3, spawn=dict(charge_dist=ah.DIST_UNIFORM, charge_b=0.9, awake=True), **kw)
    o, g = pair(oracle, gpu, cells, **kw)
    for t in range(3):
        o.step(1)
        g.step(1)
        if t == 0:
            assert np.array_equal(o.field(ah.IRRADIANCE), g.field(ah.IRRADIANCE))
    a, b = o.download(), g.download()
    compare_cells(a, b, exact=("id", "death_cause"))
    flag_diff = np.count_nonzero(a["flags"] != b["flags"])
    rng_diff = np.count_nonzero(a["rng_state"] != b["rng_state"])
    assert flag_diff == 0 and rng_diff == 0, (flag_diff, rng_diff)
    for name in ("x", "y", "z", "energy", "co2_held"):
        rel_close(a[name], b[name], 1e-9, name=name)
    rel_close(a["irradiance"], b["irradiance"], 1e-5, name="irradiance")
    rel_close(o.field(ah.TEMPERATURE), g.field(ah.TEMPERATURE), 1e-6, name="T")
    rel_close(o.field(ah.CO2), g.field(ah.CO2), 1e-5, name="CO2")


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
@pytest.mark.parametrize("n", [0, 1, 2])
def test_tick_parity_tiny_worlds(oracle, gpu, precision, n):
    """Worlds of zero, one and two awake cells (capacity equal to the count,
    one for the empty world), light, thermal and feeding on, 5 ticks: every
    per-cell pass, the hash and the field passes run over empty or nearly
    empty ranges. Integer state exact, fields and cell state as the
    lit-tick bar."""
    kw = dict(REF2D, capacity=max(n, 1), chamber_w=1e-3, chamber_h=1e-3, ambient_temp=300.0, light_enabled=1,
              light_irradiance=1.0e4, co2_init=1e-3, biology_rate=1e6, precision=precision)
    o, g = pair(oracle, gpu, **kw)
    if n:
        for w in (o, g):
            w.spawn(n, 5, charge_dist=ah.DIST_UNIFORM, charge_b=0.9, awake=True)
    o.step(5)
    g.step(5)
    a, b = o.download(), g.download()
    compare_cells(a, b)
    pos_tol = 1e-9 * 5e-4 if precision == ah.FP64_CONTRACT else FP32_POS_TOL
    for name in ("x", "y", "z"):
        rel_close(a[name], b[name], 0.0, atol=pos_tol, name=name)
    for name in ("energy", "co2_held"):
        rel_close(a[name], b[name], 1e-9, name=name)
    rel_close(o.field(ah.TEMPERATURE), g.field(ah.TEMPERATURE), 1e-6, name="T")
    rel_close(o.field(ah.CO2), g.field(ah.CO2), 1e-5, name="CO2")
    rel_close(o.field(ah.IRRADIANCE), g.field(ah.IRRADIANCE), 1e-5, name="irradiance")
    assert o.stats()["n_live"] == g.stats()["n_live"] == n
    o.close()
    g.close()


@pytest.mark.parametrize("precision", [ah.FP32_HOT, ah.FP64_CONTRACT])
def test_tick_parity_scale_workload(oracle, gpu, precision):
    """The 1M-cell scale workload itself (4 x 4 x 0.2 mm chamber, 512x512x32
    fields, 10 um hash, division and compaction on; the bench's own set-up),
    2 ticks: about two dozen divisions and two hundred deaths compacted away.
    Ids, flags, death causes and RNG streams equal the oracle's over all
    1.25M slots and the population ledger is exact. FP64 positions agree to
    1e-12 m. For the FP32 layout the oracle's float-store switch rounds
    positions and velocities to float where the device stores them, from
    the populated state on; positions then agree to FP32_POS_TOL (measured:
    2.4% of cells differ, by at most one float ulp, 2.3e-10 m). Without the
    switch the oracle keeps double positions, and the piles of deeply
    overlapping cells carry the float roundings up to 2.8e-8 m in 2 ticks."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    import bench
    cells = 1_000_000
    o = ah.World(oracle, bench.workload(oracle, cells, precision))
    g = ah.World(gpu, bench.workload(gpu, cells, precision))
    fp32 = precision == ah.FP32_HOT
    for w in (o, g):
        bench.populate(w, cells)
    if fp32:
        c = o.download()
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            c[k] = c[k].astype(np.float32).astype(np.float64)
        o.upload(c)
    oracle.lib.oracle_world_set_float_store(o.h, int(fp32))
    o.step(2)
    g.step(2)
    a, b = o.download(), g.download()
    compare_cells(a, b)
    if precision == ah.FP64_CONTRACT:
        for name in ("x", "y", "z"):
            rel_close(a[name], b[name], 0.0, atol=1e-12, name=name)
    else:
        d = np.max([np.abs(a[k] - b[k]) for k in ("x", "y", "z")], axis=0)
        far = np.count_nonzero(d > FP32_POS_TOL)
        print(f"fp32 store: {far} cells beyond {FP32_POS_TOL}, max {d.max():.3e}, "
              f"{np.count_nonzero(d)} differ", flush=True)
        assert far == 0, (far, d.max())
    for name in ("energy", "co2_held"):
        rel_close(a[name], b[name], 1e-6, name=name)
    so, sg = o.stats(), g.stats()
    for k in ("n_live", "n_dead", "n_awake", "divisions_this_window", "deaths_this_window"):
        assert so[k] == sg[k], (k, so[k], sg[k])
    assert so["divisions_this_window"] > 0 and so["deaths_this_window"] > 0
    assert o.counters().next_id == g.counters().next_id
    o.close()
    g.close()


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_crowded_feed_walk_full_rows(oracle, gpu, precision):
    """A crowded world (441 cells in 300 hash buckets of 20 um) takes the
    feed walk that gives each cell a 16-lane group, one lane per neighbour
    row. Forty cells sit upstream of cell 0 along the light, 6-9.5 um off its
    light line and inside its own (dy, dz) = (0, 0) row, so that row holds
    more survivors than the group keeps in LDS (24) and is walked again in
    order by the group's first lane. None of them, and no other cell, lies in
    cell 0's column of the 6.25 x 7.5 um irradiance voxels, so its far-field
    sample stays positive and its irradiance is the product of the forty
    factors. Irradiance must agree with the oracle to two float ulps per
    cell, as in the screen test."""
    kw = dict(capacity=500, chamber_w=200e-6, chamber_h=200e-6, chamber_d=60e-6, hash_cell_size=2e-5,
              physics_rate=0.1, thermal_enabled=0, light_enabled=1, light_irradiance=1e4, light_dir_x=1.0,
              light_dir_y=0.0, precision=precision,
              grid={0: (32, 32, 8), 1: (32, 32, 8), 2: (-1, 0, 0), 3: (32, 32, 8)})
    n, k = 441, 40
    cells = oracle_cells(oracle, n, 5, spawn=dict(awake=True, charge_dist=ah.DIST_UNIFORM, charge_b=0.9), **kw)
    rng = np.random.default_rng(11)
    # cell 0 in hash bucket (5, 5, 1) and irradiance column y [6.25, 12.5) um,
    # z [0, 7.5) um
    a0 = np.array([10e-6, 10e-6, 3.75e-6])
    off = []
    while len(off) < k:
        dy = rng.choice([-1.0, 1.0]) * rng.uniform(4.5e-6, 9e-6)
        dz = rng.uniform(-6e-6, 6e-6)
        if 6e-6 <= np.hypot(dy, dz) < 9.5e-6:
            off.append((dy, dz))
    off = np.array(off)
    cluster = np.stack([rng.uniform(-18e-6, 8e-6, k), a0[1] + off[:, 0], a0[2] + off[:, 1]], axis=1)
    pos = np.stack([cells["x"], cells["y"], cells["z"]], axis=1)
    pos[0] = a0
    pos[1:k + 1] = cluster
    rest = pos[k + 1:]
    in_col = (rest[:, 1] >= 6e-6) & (rest[:, 1] < 13e-6) & (rest[:, 2] >= -0.5e-6) & (rest[:, 2] < 8e-6)
    rest[in_col, 1] -= 30e-6
    if precision == ah.FP32_HOT:
        pos = pos.astype(np.float32).astype(np.float64)
    # every cluster cell is upstream of cell 0, within 2a of its light line,
    # in its bucket row (x buckets 4..6, y bucket 5, z bucket 1) and outside
    # its irradiance column
    d = pos[0] - pos[1:k + 1]
    assert np.all(d[:, 0] > 0.0) and np.all(np.hypot(d[:, 1], d[:, 2]) < TWO_A)
    gb = np.floor((pos[1:k + 1] - np.array([-100e-6, -100e-6, -30e-6])) / 2e-5).astype(int)
    assert np.all((gb[:, 0] >= 4) & (gb[:, 0] <= 6) & (gb[:, 1] == 5) & (gb[:, 2] == 1))
    assert np.all((pos[1:, 1] < 6.25e-6) | (pos[1:, 1] >= 12.5e-6) | (pos[1:, 2] < 0.0) | (pos[1:, 2] >= 7.5e-6))
    for j, name in enumerate(("x", "y", "z")):
        cells[name] = pos[:, j].copy()
    o, g = pair(oracle, gpu, cells, **kw)
    with o, g:
        o.step(1)
        g.step(1)
        a, b = o.download()["irradiance"], g.download()["irradiance"]
    # forty factors of 0.5-0.97 each: far below what any one neighbour leaves
    assert 0.0 < a[0] < 1e-2 * a.max(), "cell 0 should sit deep in the cluster's shadow"
    ulp = np.spacing(np.abs(a).astype(np.float32)).astype(np.float64)
    diff = np.abs(a.astype(np.float64) - b.astype(np.float64))
    bad = np.nonzero(diff > 2 * ulp)[0]
    assert bad.size == 0, f"irradiance differs at cells {bad[:10]}: {a[bad[:5]]} vs {b[bad[:5]]}"


def test_rear_cell_irradiance_exactly_zero(gpu, oracle):
    """T13 (tests/physics/test_emission.cu:73-133): collinear pair along the
    light, 15 um apart: the rear cell's irradiance and energy are exactly 0."""
    for b in (gpu, oracle):
        kw = dict(capacity=2, thermal_noise=0, thermal_enabled=0, contact_enabled=0, adhesion_enabled=0,
                  light_enabled=1, light_irradiance=1.0e4, light_dir_x=1.0, light_dir_y=0.0)
        with ah.World(b, **kw) as w:
            w.spawn(2, 1)
            c = w.download()
            c["x"][:] = [-7.5e-6, 7.5e-6]
            c["y"][:] = 0.0
            c["z"][:] = 0.0
            c["energy"][:] = 0.0
            w.upload(c)
            w.step(1)
            c = w.download()
            assert c["irradiance"][0] > 0.0 and c["energy"][0] > 0.0
            assert c["irradiance"][1] == 0.0 and c["energy"][1] == 0.0


def test_lifecycle_division_and_compaction(oracle, gpu):
    """Division slots/ids from an exclusive scan and stable compaction are
    bit-exact (T22/T22b, tests/physics/test_lifecycle.cu:194-280)."""
    kw = dict(REF2D, capacity=6000, compaction_enabled=1, co2_init=1e-3, biology_rate=1e6)
    cells = oracle_cells(oracle, 3000, 9, spawn=dict(charge_dist=ah.DIST_UNIFORM, charge_b=1.0), **kw)
    rng = np.random.default_rng(1)
    quota = 2.1e-14
    cells["co2_held"] = rng.uniform(0.9, 1.1, 3000) * quota
    dead = rng.random(3000) < 0.1
    cells["flags"][dead] &= ~np.uint32(ah.FLAG_ALIVE)
    for precision in (ah.FP64_CONTRACT, ah.FP32_HOT):
        c = dict(cells)
        if precision == ah.FP32_HOT:
            for k in ("x", "y", "z", "vx", "vy", "vz"):
                c[k] = c[k].astype(np.float32).astype(np.float64)
        o, g = pair(oracle, gpu, c, precision=precision, **kw)
        o.stage("lifecycle", 1e-3)
        g.stage("lifecycle", 1e-3)
        co, cg = o.counters(), g.counters()
        assert (co.count, co.next_id, co.divisions_total, co.deaths_total) == \
               (cg.count, cg.next_id, cg.divisions_total, cg.deaths_total)
        assert co.count < 3000 + 1500
        a, b = o.download(), g.download()
        compare_cells(a, b)
        # FP32 storage rounds each daughter coordinate to a float: half an
        # ulp at 2 mm is 1.2e-10 m.
        tol = 0.0 if precision == ah.FP64_CONTRACT else 1.25e-10
        for name in ("x", "y", "z"):
            rel_close(a[name], b[name], 1e-12, atol=tol, name=name)
        for name in ("energy", "co2_held"):
            rel_close(a[name], b[name], 1e-12, name=name)
        o.close()
        g.close()


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_tick_parity_co2_rationing(oracle, gpu, precision):
    """Whole ticks with a crowded, CO2-starved culture: most voxels hold
    several cells, so demand is rationed (lifecycle.cu:62-86), uptakes share
    accumulators and the flush sums them; division and compaction run every
    tick. Counters, slots, ids and RNG state are bit-exact and the CO2 field
    matches the oracle's bit for bit."""
    kw = dict(REF2D, capacity=6000, compaction_enabled=1, co2_init=2e-4, biology_rate=1e6, precision=precision)
    cells = oracle_cells(oracle, 3000, 21, spawn=dict(placement=ah.PLACE_GAUSSIAN, place_radius=60e-6,
                                                      charge_dist=ah.DIST_UNIFORM, charge_b=0.9), **kw)
    rng = np.random.default_rng(5)
    quota = 2.1e-14
    cells["co2_held"] = rng.uniform(0.0, 1.0, 3000) * quota
    # a tenth of the culture is a few ticks' uptake short of dividing
    near = rng.random(3000) < 0.1
    cells["co2_held"][near] = quota - rng.uniform(0.0, 4.0, near.sum()) * 6e-20
    dead = rng.random(3000) < 0.05
    cells["flags"][dead] &= ~np.uint32(ah.FLAG_ALIVE)
    if precision == ah.FP32_HOT:
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            cells[k] = cells[k].astype(np.float32).astype(np.float64)
    o, g = pair(oracle, gpu, cells, **kw)
    with o, g:
        (nx, ny, nz), _ = o.field_dims(ah.CO2)
        occupied = np.unique(np.floor((cells["x"] + 0.5 * o.desc.chamber_w) / (o.desc.chamber_w / nx)) * ny +
                             np.floor((cells["y"] + 0.5 * o.desc.chamber_h) / (o.desc.chamber_h / ny)))
        assert occupied.size < 0.5 * 3000   # most voxels are shared
        for _ in range(4):
            o.step(1)
            g.step(1)
            co, cg = o.counters(), g.counters()
            assert (co.count, co.next_id, co.divisions_total, co.deaths_total) == \
                   (cg.count, cg.next_id, cg.divisions_total, cg.deaths_total)
            assert np.array_equal(o.field(ah.CO2), g.field(ah.CO2))
        assert co.divisions_total > 0 and co.deaths_total > 0
        a, b = o.download(), g.download()
        compare_cells(a, b)
        tol = 0.0 if precision == ah.FP64_CONTRACT else FP32_POS_TOL
        for name in ("x", "y", "z"):
            rel_close(a[name], b[name], 1e-12, atol=tol, name=name)
        rel_close(a["co2_held"], b["co2_held"], 1e-12, name="co2_held")
        rel_close(a["energy"], b["energy"], 1e-12, name="energy")


def test_replay_and_graph_determinism(gpu, oracle):
    """Same inputs -> same state hash, eager or graph-replayed
    (tools/headless.cpp:348-381)."""
    kw = dict(REF2D, capacity=3000, compaction_enabled=1, co2_init=1e-3, biology_rate=1e6, light_enabled=1,
              light_irradiance=1e3, ambient_temp=330.0)
    cells = oracle_cells(oracle, 2500, 5, spawn=dict(charge_dist=ah.DIST_UNIFORM, charge_b=0.99, awake=True), **kw)
    hashes = []
    for mode in ("graph", "graph", "eager"):
        with ah.World(gpu, **kw) as w:
            w.upload(cells)
            if mode == "eager":
                w.profile(True)
            w.step(40)
            hashes.append(w.state_hash())
    assert hashes[0] == hashes[1] == hashes[2]


def test_graph_phases_odd_tick_counts(gpu, oracle):
    """A world whose CO2 field runs an odd number of substeps per tick (the
    1M-cell workload's voxel spacing: 7.8 x 7.8 x 6.25 um, dt 1e-5 s) ends
    each tick with that field's buffers swapped, so its graphs are kept per
    buffer arrangement. Stepping in odd counts, with eager ticks between
    replays, must give the state hash of the same ticks run eagerly."""
    kw = dict(capacity=3000, chamber_w=500e-6, chamber_h=500e-6, chamber_d=2e-4, hash_cell_size=1e-5,
              physics_rate=0.1, biology_rate=1e5, co2_init=1.0, division_offset=0.5e-6, compaction_enabled=1,
              grid={0: (64, 64, 32), 1: (64, 64, 32), 2: (-1, 0, 0), 3: (16, 16, 1)})
    cells = oracle_cells(oracle, 2500, 3, spawn=dict(charge_dist=ah.DIST_LOG_UNIFORM, charge_a=1e-3, charge_b=0.2),
                         **kw)
    cells["flags"][::10] |= ah.FLAG_AWAKE
    plans = {"eager": [("e", 23)], "graph": [("g", 23)],
             "mixed": [("g", 5), ("g", 1), ("e", 1), ("g", 3), ("e", 2), ("g", 1), ("g", 10)]}
    hashes = {}
    for name, plan in plans.items():
        with ah.World(gpu, **kw) as w:
            w.upload(cells)
            for mode, n in plan:
                w.profile(mode == "e")
                if mode == "e":
                    w.profile_reset()
                w.step(n)
                if mode == "e" and name == "eager":
                    kern = w.profile_read()
                    launches = kern["ftcs_CO2"][0]
                    assert launches % n == 0 and (launches // n) % 2 == 1, kern["ftcs_CO2"]
            w.profile(False)
            assert w.counters().tick == 23
            hashes[name] = w.state_hash()
    assert hashes["graph"] == hashes["eager"] and hashes["mixed"] == hashes["eager"], hashes


def test_stats_ledger(gpu, oracle):
    """world_stats' fixed-point ledger equals the host sum (T23,
    tests/physics/test_stats.cu:60-76)."""
    kw = dict(REF2D, capacity=5000)
    cells = oracle_cells(oracle, 5000, 2, spawn=dict(charge_dist=ah.DIST_UNIFORM, charge_b=1.0), **kw)
    o, g = pair(oracle, gpu, cells, **kw)
    so, sg = o.stats(), g.stats()
    for k in ("n_live", "n_dead", "n_awake", "n_dormant", "total_energy_j", "mean_temp_medium_k",
              "max_temp_medium_k", "co2_total_kg", "n2_total_kg", "mean_temp_cell_k"):
        assert so[k] == sg[k], k
    assert abs(sg["total_energy_j"] - cells["energy"].sum()) <= 1e-9 * cells["energy"].sum() + 5000 * 1e-6


@pytest.mark.parametrize("precision", [ah.FP64_CONTRACT, ah.FP32_HOT])
def test_tick_parity_real_shape(oracle, gpu, precision):
    """The plain-method real-shape workload: 25,000 awake cells with
    log-uniform charge in a 550 x 550 x 200 um box, 256 x 256 x 64 temperature
    and nutrient grids with a 5 K, sigma 40 um hot spot at the centre, light
    along +x at 1e4 W/m^2 swept over a 256 x 256 x 64 grid, dt = 1e-4 s,
    3 ticks (thermostat substeps, occlusion sweep, exact near-field shadow,
    feeding and taxis every tick). Ids, flags, death causes and RNG streams
    are exact. FP64 layout: positions and energies within 1e-9 relative.
    FP32 layout: the oracle rounds positions and velocities to float where
    the device stores them, and positions agree within FP32_POS_TOL."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    import bench
    o = ah.World(oracle, bench.workload_c2(oracle, precision))
    g = ah.World(gpu, bench.workload_c2(gpu, precision))
    bench.populate_c2(o)
    cells = o.download()
    fp32 = precision == ah.FP32_HOT
    if fp32:
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            cells[k] = cells[k].astype(np.float32).astype(np.float64)
        o.upload(cells)
    g.upload(cells)
    g.set_field(ah.TEMPERATURE, o.field(ah.TEMPERATURE))
    oracle.lib.oracle_world_set_float_store(o.h, int(fp32))
    o.step(1)
    g.step(1)
    # identical positions going into the first sweep: the swept field is exact
    assert np.array_equal(o.field(ah.IRRADIANCE), g.field(ah.IRRADIANCE))
    o.step(2)
    g.step(2)
    a, b = o.download(), g.download()
    compare_cells(a, b)
    if fp32:
        for name in ("x", "y", "z"):
            rel_close(a[name], b[name], 0.0, atol=FP32_POS_TOL, name=name)
        for name in ("energy", "co2_held"):
            rel_close(a[name], b[name], 1e-6, name=name)
        # A cell's irradiance is the far field times the exact shadow of each
        # upstream neighbour, whose overlap fraction changes by at most
        # 2 / (pi a) per metre of perpendicular offset. Positions within
        # FP32_POS_TOL then bound the difference by source x 2/(pi a) x
        # FP32_POS_TOL per shading neighbour; at this density a cell has about
        # six upstream neighbours within 2a of its light line, so eight bound it.
        a_cell = 5e-6
        irr_tol = 8 * 1e4 * (2.0 / (np.pi * a_cell)) * FP32_POS_TOL
        rel_close(a["irradiance"], b["irradiance"], 0.0, atol=irr_tol, name="irradiance")
    else:
        for name in ("x", "y", "z", "energy", "co2_held"):
            rel_close(a[name], b[name], 1e-9, name=name)
        rel_close(a["irradiance"], b["irradiance"], 1e-5, name="irradiance")
    rel_close(o.field(ah.TEMPERATURE), g.field(ah.TEMPERATURE), 1e-6, name="T")
    rel_close(o.field(ah.CO2), g.field(ah.CO2), 1e-5, name="CO2")
    so, sg = o.stats(), g.stats()
    for k in ("n_live", "n_dead", "n_awake"):
        assert so[k] == sg[k], (k, so[k], sg[k])
    assert so["n_awake"] == 25_000
    o.close()
    g.close()
