"""Algorithmic bytes per kernel launch, for the HBM-roofline figures.

"Algorithmic" means compulsory traffic: every per-cell or per-voxel array a
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
        # + the voxel and concentration each live cell books
        "integrate_demand": integ + 4 * n + alive * 16,
        "co2_demand": alive * (20 + 3 * p) + 4 * n,
        # voxel, flags, t_local, lslot per cell; concentration, demand and
        # the store per live cell
        "uptake_mark": 16 * n + alive * 36,
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

