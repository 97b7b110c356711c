"""The oracle's predation step (src/sim/predation.cu) on the CPU, pinned by
what its rules fix rather than by restating them:

* the crawl: with noise off, one tick brings a predator to the canon crawl
  speed along its stored heading (TAU_CRAWL_THRUST / drag = TAU_CRAWL_SPEED,
  tests/physics/test_predation.cu:64-66), and with no prey it tumbles every
  tick (the rule's signal never rises);
* the engulf: an overlapping live cell dies Predated, its store goes per the
  disposition, the predator digests for TAU_DIGEST_TIME of biology time and
  then banks TAU_BIOMASS_YIELD of the prey's biomass; of two claimants the
  lower id wins;
* nitrogen: deaths over many predators follow 1 - exp(-hazard * rate *
  dt_bio) within five standard deviations, and tolerance raises the
  threshold linearly;
* division at twice the dry biomass: halves, the daughter one id up,
  generation + 1, the pair two radii apart, and stable compaction;
* the reference's T30 gates: bit-reproducible, seed-sensitive, fewer live
  cells, every predator inside the chamber.
The helpers themselves are checked bit for bit against the reference's code
in test_oracle_golden.py.
"""
import math

import numpy as np
import pytest

from conftest import ah

CRAWL_SPEED = 5.0000000000000004e-06
TAU_RADIUS = 2.0000000000000002e-05
DIGEST_TIME = 120.0
YIELD = 0.59999999999999998
SMALL = dict(chamber_w=3e-4, chamber_h=3e-4, chamber_d=6e-5, emission_enabled=0, taxis_enabled=0,
             thermal_enabled=0, grid={0: (32, 32, 1), 1: (16, 16, 1), 2: (16, 16, 1), 3: (32, 32, 1)})


def world(oracle, **kw):
    return ah.World(oracle, **dict(SMALL, **kw))


def test_crawl_at_the_canon_speed(oracle):
    with world(oracle, capacity=10, thermal_noise=0) as w:
        w.tau_spawn(5, 3)
        for _ in range(3):
            w.step(1)
            t = w.tau_download()
            v = np.stack([t["vx"], t["vy"], t["vz"]], 1)
            d = np.stack([t["dir_x"], t["dir_y"], t["dir_z"]], 1).astype(np.float64)
            speed = np.linalg.norm(v, axis=1)
            # x = gamma dt / m ~ 11: the previous heading's velocity decays by e^-x
            assert np.all(np.abs(speed / CRAWL_SPEED - 1.0) < 5e-5), speed
            cos = np.sum(v * d, 1) / speed / np.linalg.norm(d, axis=1)
            assert np.all(cos > 1.0 - 1e-8)
            assert np.all(t["run_timer"] == 0.0)   # no prey: a tumble every tick


def place(oracle_world, tau_xyz, cell_xyz, cell_biomass=2.1e-14):
    """One dormant live cell per position, predators at their positions."""
    w = oracle_world
    w.spawn(len(cell_xyz), 9)
    c = w.download()
    for k, (x, y, z) in enumerate(cell_xyz):
        c["x"][k], c["y"][k], c["z"][k] = x, y, z
    c["biomass"][:] = cell_biomass
    c["energy"][:] = 7.5e5
    w.upload(c)
    w.tau_spawn(len(tau_xyz), 4)
    t = w.tau_download()
    for k, (x, y, z) in enumerate(tau_xyz):
        t["x"][k], t["y"][k], t["z"][k] = x, y, z
    w.tau_upload(t)


@pytest.mark.parametrize("disposition", [ah.STORE_VOID, ah.STORE_RETAIN])
def test_engulf_then_digest(oracle, disposition):
    # dt_bio = 1e-3 s * 1e5 = 100 s: digestion ends two ticks after the engulf
    with world(oracle, capacity=10, biology_rate=1e5, lifecycle_enabled=0, store_disposition=disposition) as w:
        place(w, [(0.0, 0.0, 0.0)], [(1e-5, 0.0, 0.0)], cell_biomass=3.0e-14)
        w.step(1)
        c, t = w.download(), w.tau_download()
        assert not (c["flags"][0] & ah.FLAG_ALIVE) and c["death_cause"][0] == ah.DEATH_PREDATED
        assert c["energy"][0] == (7.5e5 if disposition == ah.STORE_RETAIN else 0.0)
        assert t["digest_timer"][0] == np.float32(DIGEST_TIME) and t["prey_biomass"][0] == 3.0e-14
        assert t["target"][0] == 0
        b0 = t["biomass"][0]
        w.step(1)
        t = w.tau_download()
        assert t["digest_timer"][0] == np.float32(np.float32(DIGEST_TIME) - np.float32(100.0))
        w.step(1)
        t = w.tau_download()
        assert t["digest_timer"][0] == 0.0 and t["prey_biomass"][0] == 0.0
        assert t["biomass"][0] == b0 + 3.0e-14 * YIELD


def test_lowest_id_claims_the_cell(oracle):
    with world(oracle, capacity=10, lifecycle_enabled=0) as w:
        place(w, [(2e-5, 0.0, 0.0), (-2e-5, 0.0, 0.0)], [(0.0, 0.0, 0.0)])
        w.step(1)
        t = w.tau_download()
        assert list(t["id"]) == [1, 2]
        assert list(t["target"]) == [0, 0]           # both claimed slot 0
        assert t["digest_timer"][0] > 0 and t["digest_timer"][1] == 0.0


def test_nitrogen_deaths_follow_the_hazard(oracle):
    n = 4000
    rate, n2 = 1e5, 0.015
    with world(oracle, capacity=10, tau_capacity=n, biology_rate=rate) as w:
        w.tau_spawn(n, 6)
        t = w.tau_download()
        t["tolerance"][: n // 2] = 1.0               # threshold 0.02 > 0.015: immune
        w.tau_upload(t)
        w.fill(ah.N2, n2)
        w.step(1)
        alive = (w.tau_download()["flags"] & ah.FLAG_ALIVE) != 0
        assert alive[: n // 2].all()
        p = -math.expm1(-(n2 - 0.01) * 0.05 * (1e-3 * rate))
        dead = int((~alive[n // 2:]).sum())
        mu, sd = p * (n // 2), math.sqrt(p * (1 - p) * (n // 2))
        assert abs(dead - mu) < 5 * sd, (dead, mu, sd)
        st = w.stats()
        assert st["n_taumoeba"] == int(alive.sum())
        assert abs(st["mean_tau_tolerance"] - (n // 2) / alive.sum()) < 1e-6


@pytest.mark.parametrize("compact", [0, 1])
def test_division_and_compaction(oracle, compact):
    with world(oracle, capacity=10, tau_capacity=16, lifecycle_enabled=0, tau_compaction_enabled=compact) as w:
        w.tau_spawn(3, 8)
        t = w.tau_download()
        t["biomass"][1] = 2.4e-13                    # >= 2 x TAU_MASS_DRY: divides
        t["flags"][0] &= ~np.uint32(ah.FLAG_ALIVE)    # a corpse in slot 0
        w.tau_upload(t)
        n0, _, next0 = w.tau_counters()
        w.step(1)
        t = w.tau_download()
        n1, _, next1 = w.tau_counters()
        assert next1 == next0 + 1
        if compact:
            assert n1 == 3 and list(t["id"]) == [2, 3, 4]   # stable: survivors, then the daughter
            par, dau = 0, 2
        else:
            assert n1 == n0 + 1 and list(t["id"]) == [1, 2, 3, 4]
            par, dau = 1, 3
        assert t["biomass"][par] == t["biomass"][dau] == 1.2e-13
        assert t["generation"][dau] == 1 and t["generation"][par] == 0
        assert 0.0 <= t["tolerance"][dau] <= 1.0 and abs(t["tolerance"][dau] - t["tolerance"][par]) < 0.15
        sep = math.dist((t["x"][par], t["y"][par], t["z"][par]), (t["x"][dau], t["y"][dau], t["z"][dau]))
        # along a unit axis, unless the chamber wall clamps one of them (60 um deep)
        assert sep <= 2 * TAU_RADIUS * (1 + 1e-12)
        assert t["target"][dau] == -1 and t["digest_timer"][dau] == 0.0


def crowded(oracle, seed):
    w = world(oracle, capacity=2600, tau_capacity=64, co2_init=0.0, biology_rate=1e5, seed=seed)
    w.spawn(2000, seed, charge_dist=ah.DIST_CONSTANT, charge_a=0.03)
    w.tau_spawn(30, seed)
    return w


def test_t30_reproducible_contained_and_eating(oracle):
    """T30.2/T30.3 (tests/physics/test_predation.cu:83-150) on a crowded
    chamber: two runs agree bit for bit, another seed differs, predators
    eat, and every predator stays inside the chamber."""
    hashes = []
    for seed in (11, 11, 12):
        with crowded(oracle, seed) as w:
            live0 = w.stats()["n_live"]
            w.step(150)
            hashes.append(w.state_hash())
            st = w.stats()
            assert st["n_live"] < live0
            t = w.tau_download()
            a = TAU_RADIUS
            for k, half in (("x", 1.5e-4), ("y", 1.5e-4), ("z", 3e-5)):
                assert np.all(np.abs(t[k]) <= half + a)
            c = w.download()
            assert np.count_nonzero(c["death_cause"] == ah.DEATH_PREDATED) == live0 - st["n_live"]
    assert hashes[0] == hashes[1] != hashes[2]
