"""Algorithmic bytes per kernel launch, for the HBM-roofline figures.

Two counts live here.

`survey_kernel_bytes` is the basis of every reported roofline fraction: the
per-stage compulsory traffic of the plain step, frozen in STAGE_BYTES (bytes
per cell, per awake or live cell, per bucket, per grid point), shared out over
the kernels that implement each stage by KERNEL_STAGES. The table is fixed:
it does not follow this build's data layout, so a layout that moves more
bytes than the stage needs shows up as a lower fraction, and one that skips
traffic the reference's design pays for gets it credited as the reference
would need it.

`kernel_bytes` is the builder's own, finer count of what each kernel of this
layout must touch; it is printed beside the basis, never in its place.

For the builder's count, "algorithmic" means compulsory traffic: every per-cell or per-voxel array a
kernel must touch is counted once per read and once per write, at its stored
width. Gathers from a field grid count one element; neighbour records and
bucket starts reached by the contact and shadow walks count zero (they are
re-reads of data another thread of the same pass fetched). Per-cell sizes
depend on the position width `p` (4 for the FP32 layout, 8 for FP64) and the
neighbour-record size `r` (16 / 32 B).

`counts` carries the population figures of the interval measured: n (cells),
awake, alive, births, dead (per tick), compacting (fraction of ticks that
compacted), plus grid sizes: pts[field] and buckets.
"""
from __future__ import annotations

# bytes per cell of the whole SoA in each layout (contracts/cell_store_v1.h
# order; positions and velocities at the layout's width)
SOA_BYTES = {4: 117, 8: 141}


def kernel_bytes(name: str, c: dict) -> float:
    n, p = c["n"], c["p"]
    r = 16 if p == 4 else 32
    awake, alive = c.get("awake", 0), c.get("alive", n)
    pts = c["pts"]
    k_fields = c.get("sampled_fields", 2)
    S = SOA_BYTES[p]
    births, dead = c.get("births", 0), c.get("dead", 0)
    compacting = c.get("compacting", 0)
    resort_frac = c.get("resort_frac", 0.0)
    # per-cell bytes of the integrator: flags, t_local, biomass, energy, the
    # force (via spos), position and velocity, id, RNG state and age in;
    # RNG state, flags, position, velocity and age out
    integ = n * (88 + 12 * p)
    table = {
        "hash_key": n * (16 + 3 * p),
        "hash_bucket_scan": 8 * (c["buckets"] + 1) + 4 * min(n, c["buckets"]),
        "hash_scatter": n * (16 + 3 * p + r),
        "hash_rank": n * (2 * r + 16),
        # a re-sort moves the whole SoA twice (gather, copy back), every
        # `resort_frac` of the ticks
        "resort_gather": resort_frac * n * (2 * S + 4),
        "resort_copyback": resort_frac * n * 2 * S,
        "thermal": n * (12 + 3 * p) + awake * 24,
        "deposit_apply": 4 * n + awake * (3 * p + 16),
        "ftcs_T": 8 * pts.get(0, 0),
        "ftcs_CO2": 8 * pts.get(1, 0),
        "ftcs_N2": 8 * pts.get(2, 0),
        "taxis": 4 * n + 4 * max(alive - awake, 0) + 28 * awake,
        "feed_taxis": 4 * n + 24 * alive + 28 * awake,
        "feed": 4 * n + 24 * alive,
        # lslot, flags, position, biomass, energy, emit power, spos in; the
        # local field samples (one value per sampled field) and the force out
        "sample_base": n * (4 + 4 + 3 * p + 8 + 8 + 4 + 4 + 4 * k_fields + 12 + 24) + awake * 12,
        # own record, the force read and written; neighbour records and
        # bucket starts are re-reads of data other lanes fetched
        "contact_walk": n * (r + 48),
        # sample_base with feeding and/or taxis in the same pass: both
        # kernels' bytes, the flags read once
        "taxis_sample": 0.0,
        "feed_sample": 0.0,
        "feed_taxis_sample": 0.0,
        "integrate": integ,
        # + the voxel and concentration each live cell books (its voxel's
        # 4-byte count of booking cells included)
        "integrate_demand": integ + 4 * n + alive * 12,
        "co2_demand": alive * (16 + 3 * p) + 4 * n,
        # voxel, flags, t_local, lslot per cell; concentration, the voxel's
        # booking count and the store per live cell
        "uptake_mark": 16 * n + alive * 32,
        "life_scan": 8 * n * ((1 if births > 0 else 0) + compacting),
        # the slot of every entry (flush and renumbering) and each daughter
        "life_apply": 8 * n + births * 2 * S,
        "life_finalize": 0,
        "compact_mark": 8 * n,
        "occlusion_stamp": n * (4 + 3 * p + 8),
        "light_sweep": 12 * pts.get(3, 0),
        "feed_walk": alive * (r + 8 + 4),
        "spawn": n * S,
    }
    for fused, part in (("taxis_sample", "taxis"), ("feed_sample", "feed"), ("feed_taxis_sample", "feed_taxis")):
        table[fused] = table["sample_base"] + table[part] - 4 * n
    return float(table.get(name, 0.0))



# ---------------------------------------------------------------------------
# The frozen basis: compulsory bytes per unit of each stage of the plain step
# (hot FP32 cell layout, FP32 fields). Units: "cell" = every slot of the
# population, "awake" / "alive" = those cells only, "bucket" = hash bucket,
# "point" = grid point of the field named, per launch of the stage.
# ---------------------------------------------------------------------------
STAGE_BYTES = {
    "hash_keys": (24, "cell"),
    "sort_pass": (16, "cell"),          # per radix pass; passes = ceil(key bits / 11)
    "sort_tail": (4, "cell"),
    "ranges": (4, "cell"),
    "buckets": (8, "bucket"),           # clear + range write
    "field_sample": (24, "cell"),
    "forces": (52, "cell"),
    "integrate_pcg": (108, "cell"),
    "integrate_philox": (92, "cell"),
    "thermal": (28, "cell"),            # per temperature substep
    "light": (44, "cell"),              # lit worlds
    "taxis": (84, "awake"),
    "lifecycle": (72, "cell"),
    "compaction": (12 + 2 * 93, "cell"),   # ticks that compact
    "ftcs": (8, "point"),               # per substep, per field
    "fold": (8, "point"),               # deposit fold: per T substep, per tick for the nutrient
    "irradiance": (20, "point"),        # per tick when lit
}

# Kernel -> [(stage, share of that stage's bytes, grid the points count on)].
# Shares of one stage over its kernels add to 1.
KERNEL_STAGES = {
    "hash_key": [("hash_keys", 1.0, None)],
    "hash_bucket_scan": [("buckets", 1.0, None)],
    "hash_scatter": [("sort_pass", 1.0, None), ("sort_tail", 0.5, None)],
    "hash_rank": [("sort_pass", 1.0, None), ("sort_tail", 0.5, None), ("ranges", 1.0, None)],
    "thermal": [("thermal", 1.0, None)],
    "deposit_apply": [("fold", 1.0, 0)],
    "ftcs_T": [("ftcs", 1.0, 0)],
    "ftcs_CO2": [("ftcs", 1.0, 1)],
    "ftcs_N2": [("ftcs", 1.0, 2)],
    "sample_base": [("field_sample", 1.0, None)],
    "taxis_sample": [("field_sample", 1.0, None), ("taxis", 1.0, None)],
    "feed_sample": [("field_sample", 1.0, None), ("light", 0.25, None)],
    "feed_taxis_sample": [("field_sample", 1.0, None), ("light", 0.25, None), ("taxis", 1.0, None)],
    "feed": [("light", 0.25, None)],
    "taxis": [("taxis", 1.0, None)],
    "feed_taxis": [("light", 0.25, None), ("taxis", 1.0, None)],
    "contact_walk": [("forces", 1.0, None)],
    "integrate": [("integrate", 1.0, None)],
    "integrate_demand": [("integrate", 1.0, None), ("lifecycle", 0.25, None)],
    "uptake_mark": [("lifecycle", 0.5, None), ("fold", 1.0, 1)],
    "life_scan": [("lifecycle", 0.1, None)],
    "life_apply": [("lifecycle", 0.15, None), ("compaction", 1.0, None)],
    "occlusion_stamp": [("light", 0.25, None)],
    "light_sweep": [("irradiance", 1.0, 3)],
    "feed_walk": [("light", 0.25, None)],
}


def radix_passes(buckets: int) -> int:
    # keys run 0 .. buckets (the free-slot sentinel): that many bits, 11 a pass
    bits = max(1, int(buckets).bit_length())
    return -(-bits // 11)


def survey_kernel_bytes(name: str, c: dict) -> float:
    """Bytes per launch of kernel `name` on the frozen basis. `c` holds n,
    awake, alive, buckets, pts{field: points}, compacting (fraction of
    ticks), philox (bool) and, for lifecycle kernels, nothing else."""
    total = 0.0
    for stage, share, field in KERNEL_STAGES.get(name, []):
        if stage == "integrate":
            stage = "integrate_philox" if c.get("philox") else "integrate_pcg"
        per, unit = STAGE_BYTES[stage]
        if unit == "cell":
            units = c["n"]
        elif unit == "awake":
            units = c.get("awake", 0)
        elif unit == "alive":
            units = c.get("alive", c["n"])
        elif unit == "bucket":
            units = c["buckets"]
        else:
            units = c["pts"].get(field, 0)
        if stage == "sort_pass":
            units *= radix_passes(c["buckets"]) / 2.0   # the two sort kernels share the passes
        if stage == "compaction":
            units *= c.get("compacting", 0.0)
        total += share * per * units
    return float(total)
