"""Python host mirror of the hot-path C ABI (include/astro_hot.h).

The reference drives its world from C++ (src/sim/world.cuh:183-243); this
module gives the same calls to Python callers, tests and the benchmark, over
ctypes. It mirrors the reference's names and error behaviour: every entry point
returns an astro status code, and a non-zero code raises AstroError with the
world's error detail.

`load()` binds the HIP library built in-tree by build.py. It raises when the
library is missing or no GPU is visible: there is no CPU fallback.

`Binding` is generic over the symbol prefix so that a caller holding another
library with the same ABI (the CPU restatement used by the tests) can drive it
through the same wrapper.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libastro_hot.so"

ABI_VERSION = 2
STATUS = ["Ok", "OutOfMemory", "DeviceError", "InvalidArgument", "NotFound", "ParseError",
          "ContractVersionMismatch", "CapacityExceeded", "FileIoError", "Unsupported"]
OK, OUT_OF_MEMORY, DEVICE_ERROR, INVALID_ARGUMENT, NOT_FOUND, PARSE_ERROR, \
    CONTRACT_VERSION_MISMATCH, CAPACITY_EXCEEDED, FILE_IO_ERROR, UNSUPPORTED = range(10)

FP64_CONTRACT, FP32_HOT = 0, 1
RNG_PCG32, RNG_PHILOX = 0, 1
RNG_STAGE_SPAWN, RNG_STAGE_TAXIS, RNG_STAGE_INTEGRATE = 1, 2, 3
DEPOSIT_AUTO, DEPOSIT_ATOMIC, DEPOSIT_DIRECT = 0, 1, 2
BOUNDARY_REFLECTING, BOUNDARY_PERIODIC, BOUNDARY_ABSORBING = 0, 1, 2
GRAVITY_Y, GRAVITY_Z = 0, 1
CLOCK_REALTIME, CLOCK_MOTION, CLOCK_METABOLIC, CLOCK_GENERATIONAL, CLOCK_CUSTOM = range(5)
BRUSH_HEAT, BRUSH_CHILL, BRUSH_INJECT_CO2, BRUSH_INJECT_N2 = range(4)
STORE_VOID, STORE_FLASH, STORE_RETAIN = 0, 1, 2
TEMPERATURE, CO2, N2, IRRADIANCE = 0, 1, 2, 3
PLACE_UNIFORM, PLACE_GAUSSIAN, PLACE_GRID, PLACE_DISC = 0, 1, 2, 3
DIST_CONSTANT, DIST_UNIFORM, DIST_NORMAL, DIST_LOG_UNIFORM = 0, 1, 2, 3

# contract flags (contracts/cell_store_v1.h:20-28)
FLAG_OCCUPIED, FLAG_ALIVE, FLAG_AWAKE, FLAG_STUCK = 1, 2, 4, 8
DEATH_NONE, DEATH_STARVED, DEATH_PREDATED, DEATH_OVERHEATED, DEATH_CULLED = range(5)

# Cell SoA in contract-v1 order (contracts/cell_store_v1.h:42-74).
CELL_FIELDS = [
    ("id", np.uint64), ("flags", np.uint32), ("death_cause", np.uint8),
    ("x", np.float64), ("y", np.float64), ("z", np.float64),
    ("vx", np.float64), ("vy", np.float64), ("vz", np.float64),
    ("energy", np.float64), ("temp_cell", np.float32), ("emit_power", np.float32),
    ("dir_x", np.float32), ("dir_y", np.float32), ("dir_z", np.float32),
    ("biomass", np.float64), ("co2_held", np.float64), ("age_s", np.float32),
    ("rng_state", np.uint64), ("taxis_memory", np.float32), ("run_timer", np.float32),
    ("t_local", np.float32), ("irradiance", np.float32), ("co2_local", np.float32),
    ("n2_local", np.float32),
]


# TaumoebaStore arrays in the reference's order (src/sim/predation.cuh:85-104).
TAU_FIELDS = [
    ("id", np.uint64), ("flags", np.uint32), ("x", np.float64), ("y", np.float64), ("z", np.float64),
    ("vx", np.float64), ("vy", np.float64), ("vz", np.float64), ("dir_x", np.float32), ("dir_y", np.float32),
    ("dir_z", np.float32), ("biomass", np.float64), ("prey_biomass", np.float64), ("digest_timer", np.float32),
    ("tolerance", np.float32), ("generation", np.uint32), ("rng_state", np.uint64), ("density_ema", np.float32),
    ("run_timer", np.float32), ("target", np.int32),
]


class AstroError(RuntimeError):
    def __init__(self, status: int, detail: str = ""):
        name = STATUS[status] if 0 <= status < len(STATUS) else f"status {status}"
        super().__init__(f"{name}: {detail}" if detail else name)
        self.status = status


class GridDims(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32)]


class WorldDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("capacity", C.c_int32), ("seed", C.c_uint64),
        ("chamber_w", C.c_double), ("chamber_h", C.c_double), ("chamber_d", C.c_double),
        ("precision", C.c_int32), ("tau_capacity", C.c_int32), ("hash_cell_size", C.c_double),
        ("grid", GridDims * 4), ("_pad1", C.c_int32),
        ("boundary_x", C.c_uint8), ("boundary_y", C.c_uint8), ("gravity_axis", C.c_uint8),
        ("thermal_noise", C.c_uint8), ("contact_enabled", C.c_uint8), ("adhesion_enabled", C.c_uint8),
        ("thermal_enabled", C.c_uint8), ("emission_enabled", C.c_uint8), ("taxis_enabled", C.c_uint8),
        ("occlusion_exact", C.c_uint8), ("store_disposition", C.c_uint8), ("compaction_enabled", C.c_uint8),
        ("lifecycle_enabled", C.c_uint8), ("tau_compaction_enabled", C.c_uint8), ("rng_kind", C.c_uint8),
        ("_pad2", C.c_uint8 * 5),
        ("ambient_temp", C.c_double), ("co2_init", C.c_double), ("division_offset", C.c_double),
        ("physics_rate", C.c_double), ("biology_rate", C.c_double),
        ("petrova_max_power", C.c_double), ("co2_mass_per_division", C.c_double),
        ("light_dir_x", C.c_float), ("light_dir_y", C.c_float), ("light_irradiance", C.c_float),
        ("ambient_irradiance", C.c_float), ("light_enabled", C.c_uint8), ("_pad3", C.c_uint8 * 7),
        ("petrova_flash_power", C.c_double),
    ]


class SpawnParams(C.Structure):
    _fields_ = [
        ("count", C.c_int32), ("placement", C.c_int32),
        ("place_x", C.c_double), ("place_y", C.c_double), ("place_radius", C.c_double),
        ("charge_dist", C.c_int32), ("awake", C.c_int32),
        ("charge_a", C.c_double), ("charge_b", C.c_double),
    ]


class CellSoA(C.Structure):
    _fields_ = [(name, C.c_void_p) for name, _ in CELL_FIELDS]


class TauSoA(C.Structure):
    _fields_ = [(name, C.c_void_p) for name, _ in TAU_FIELDS]


class Counters(C.Structure):
    _fields_ = [
        ("count", C.c_int32), ("capacity", C.c_int32), ("next_id", C.c_uint64), ("tick", C.c_uint64),
        ("sim_time_s", C.c_double), ("divisions_total", C.c_int64), ("deaths_total", C.c_int64),
        ("dead_slots_prev", C.c_int32), ("_pad", C.c_int32), ("cell_steps", C.c_int64),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("tick", C.c_uint64), ("sim_time_s", C.c_double),
        ("n_live", C.c_int32), ("n_dead", C.c_int32), ("n_dormant", C.c_int32), ("n_awake", C.c_int32),
        ("n_taumoeba", C.c_int32), ("_pad0", C.c_int32),
        ("mean_charge", C.c_double), ("total_energy_j", C.c_double), ("mean_temp_cell_k", C.c_double),
        ("mean_temp_medium_k", C.c_double), ("max_temp_medium_k", C.c_double), ("co2_total_kg", C.c_double),
        ("n2_total_kg", C.c_double), ("mean_tau_tolerance", C.c_double),
        ("divisions_this_window", C.c_int32), ("deaths_this_window", C.c_int32),
        ("boil_events", C.c_int32), ("ignitions_this_window", C.c_int32),
        ("physics_rate", C.c_double), ("biology_rate", C.c_double),
        ("non_canon_run", C.c_uint8), ("_pad1", C.c_uint8 * 7),
    ]


assert C.sizeof(WorldDesc) == 216, C.sizeof(WorldDesc)
assert C.sizeof(SpawnParams) == 56
assert C.sizeof(Counters) == 64
assert C.sizeof(Stats) == 144


# ---- scenarios (include/astro_scenario.h) -----------------------------------
ORGANISM_ASTROPHAGE, ORGANISM_TAUMOEBA = 0, 1
STIM_HEAT, STIM_CHILL, STIM_INJECT_CO2, STIM_INJECT_N2, STIM_SET_LIGHT, STIM_SET_N2, STIM_SEED_CELLS, \
    STIM_FLASH = range(8)
METRICS = ["awake_fraction", "medium_temp_mean", "medium_temp_max", "boil_event_count", "population_live",
           "population_dead", "mean_charge", "total_energy_j", "mean_tau_tolerance", "max_tau_tolerance",
           "charge_depth_correlation", "charge_height_correlation", "rise_velocity_empty",
           "fall_velocity_full", "doubling_time_s", "impulse_per_cycle"]
OPS = ["==", "!=", "<", "<=", ">", ">=", "~="]


class LightSource(C.Structure):
    _fields_ = [("dir_x", C.c_float), ("dir_y", C.c_float), ("irradiance", C.c_float),
                ("wavelength", C.c_float), ("enabled", C.c_int32)]


class PopulationSpec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("count", C.c_int32), ("placement", C.c_int32), ("charge_dist", C.c_int32),
                ("place_x", C.c_double), ("place_y", C.c_double), ("place_radius", C.c_double),
                ("charge_a", C.c_double), ("charge_b", C.c_double), ("awake", C.c_int32), ("_pad", C.c_int32)]


class Stimulus(C.Structure):
    _fields_ = [("t0_s", C.c_double), ("t1_s", C.c_double), ("x", C.c_double), ("y", C.c_double),
                ("radius", C.c_double), ("v0", C.c_double), ("v1", C.c_double),
                ("kind", C.c_int32), ("_pad", C.c_int32)]


class AcceptCheck(C.Structure):
    _fields_ = [("metric", C.c_int32), ("op", C.c_int32), ("value", C.c_double), ("tol", C.c_double),
                ("after_s", C.c_double), ("tol_absolute", C.c_int32), ("_pad", C.c_int32)]


class ParamOverride(C.Structure):
    _fields_ = [("key", C.c_char * 48), ("value", C.c_double)]


class Scenario(C.Structure):
    _fields_ = [
        ("schema", C.c_int32), ("id", C.c_char * 64), ("title", C.c_char * 128), ("blurb", C.c_char * 256),
        ("seed", C.c_uint64), ("chamber_w", C.c_double), ("chamber_h", C.c_double), ("chamber_d", C.c_double),
        ("boundary_x", C.c_int32), ("boundary_y", C.c_int32),
        ("temp_init", C.c_double), ("co2_init", C.c_double), ("n2_init", C.c_double), ("ambient_temp", C.c_double),
        ("thermal_bc", C.c_int32), ("density_model", C.c_int32), ("store_disposition", C.c_int32),
        ("gravity_axis", C.c_int32), ("thermal_active", C.c_int32), ("compaction", C.c_int32),
        ("tau_compaction", C.c_int32), ("ambient_irradiance", C.c_float),
        ("lights", LightSource * 8), ("populations", PopulationSpec * 8), ("population_count", C.c_int32),
        ("clock", C.c_int32), ("physics_rate", C.c_double), ("biology_rate", C.c_double),
        ("tools", C.c_uint32), ("override_count", C.c_int32), ("param_overrides", ParamOverride * 32),
        ("drive", Stimulus * 16), ("drive_count", C.c_int32), ("objective_text", C.c_char * 256),
        ("accept_count", C.c_int32), ("accept", AcceptCheck * 16), ("run_duration_s", C.c_double),
    ]


class AcceptResult(C.Structure):
    _fields_ = [("metric", C.c_int32), ("op", C.c_int32), ("measured", C.c_double), ("value", C.c_double),
                ("tol", C.c_double), ("after_s", C.c_double), ("tol_absolute", C.c_int32), ("pass_", C.c_int32)]


class AssertSummary(C.Structure):
    _fields_ = [("ticks", C.c_int64), ("sim_time_s", C.c_double), ("checks", C.c_int32), ("passed", C.c_int32),
                ("all_pass", C.c_int32), ("samples", C.c_int32)]


assert C.sizeof(Scenario) == 4992 and Scenario.accept.offset == 4344 and Scenario.drive.offset == 3056
assert C.sizeof(AcceptResult) == 48 and C.sizeof(AssertSummary) == 32

# Every entry point of include/astro_hot.h: name -> (restype, argtypes).
_P = C.c_void_p
_I = C.c_int32
_D = C.c_double
SIGNATURES = {
    "philox_blocks": (_I, [C.c_void_p, _I, C.c_void_p]),
    "philox_stream": (_I, [C.c_uint64, C.c_uint64, C.c_uint64, _I, _I, C.c_void_p]),
    "world_desc_default": (None, [C.POINTER(WorldDesc)]),
    "world_create": (_I, [C.POINTER(WorldDesc), C.POINTER(_P)]),
    "world_destroy": (None, [_P]),
    "status_str": (C.c_char_p, [_I]),
    "last_error": (C.c_char_p, [_P]),
    "world_desc_get": (_I, [_P, C.POINTER(WorldDesc)]),
    "cells_spawn": (_I, [_P, C.POINTER(SpawnParams), C.c_uint64]),
    "cells_upload": (_I, [_P, C.POINTER(CellSoA), _I]),
    "cells_download": (_I, [_P, C.POINTER(CellSoA), _I, C.POINTER(_I)]),
    "world_get_counters": (_I, [_P, C.POINTER(Counters)]),
    "world_set_counters": (_I, [_P, C.POINTER(Counters)]),
    "field_dims": (_I, [_P, _I, C.POINTER(GridDims), C.POINTER(_D), C.POINTER(_D), C.POINTER(_D)]),
    "field_upload": (_I, [_P, _I, _P, C.c_size_t]),
    "field_download": (_I, [_P, _I, _P, C.c_size_t]),
    "field_deposit_download": (_I, [_P, _I, _P, C.c_size_t]),
    "world_step": (_I, [_P, _I]),
    "world_sync": (_I, [_P]),
    "world_stats": (_I, [_P, C.POINTER(Stats)]),
    "world_state_hash": (_I, [_P, C.POINTER(C.c_uint64)]),
    "world_set_clock": (_I, [_P, _D, _D]),
    "world_set_clock_preset": (_I, [_P, _I, _D, _D]),
    "world_set_light": (_I, [_P, C.c_float, C.c_float, C.c_float, C.c_float, _I]),
    "world_apply_brush": (_I, [_P, _I, _D, _D, _D, _D]),
    "field_fill": (_I, [_P, _I, C.c_float]),
    "world_set_flash": (_I, [_P, _I]),
    "world_flash_audit": (_I, [_P, C.POINTER(_D), C.POINTER(_D)]),
    "world_time": (_I, [_P, C.POINTER(C.c_uint64), C.POINTER(_D), C.POINTER(_D), C.POINTER(_D)]),
    "stage_hash_build": (_I, [_P]),
    "hash_dims": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), C.POINTER(_D)]),
    "hash_download": (_I, [_P, _P, _I, _P, _P, _I]),
    "stage_thermal": (_I, [_P, _D]),
    "stage_emission": (_I, [_P, _D]),
    "stage_taxis": (_I, [_P, _D]),
    "stage_forces": (_I, [_P]),
    "forces_download": (_I, [_P, _P, _P, _P, _I]),
    "stage_integrate": (_I, [_P, _D]),
    "stage_motion": (_I, [_P, _D]),
    "stage_slow_fields": (_I, [_P, _D]),
    "stage_diffuse": (_I, [_P, _I, _D]),
    "stage_lifecycle": (_I, [_P, _D]),
    "stage_compact": (_I, [_P]),
    "taumoeba_spawn": (_I, [_P, _I, C.c_uint64]),
    "taumoeba_upload": (_I, [_P, C.POINTER(TauSoA), _I]),
    "taumoeba_download": (_I, [_P, C.POINTER(TauSoA), _I, C.POINTER(_I)]),
    "taumoeba_counters": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(C.c_uint64)]),
    "stage_predation": (_I, [_P, _D]),
}
PRODUCT_ONLY = {
    "hash_build_noalloc": (_I, [_P, _P, _P, _P, _I, _I, _I, _I, _D, _P, _P, _P, _P]),
    "world_set_deposit_path": (_I, [_P, _I]),
    "profile_enable": (_I, [_P, _I]),
    "profile_read": (_I, [_P, _P, _I, _P, _P, _I, C.POINTER(_I)]),
    "profile_reset": (_I, [_P]),
    "world_stream": (_P, [_P]),
    "snapshot_to_bytes": (_I, [_P, C.c_char_p, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "snapshot_save": (_I, [_P, C.c_char_p, C.c_char_p]),
    "snapshot_from_bytes": (_I, [C.POINTER(WorldDesc), _P, C.c_size_t, C.POINTER(_P)]),
    "snapshot_load": (_I, [C.POINTER(WorldDesc), C.c_char_p, C.POINTER(_P)]),
    "ensemble_create": (_I, [C.POINTER(_P), _I, C.POINTER(_P)]),
    "ensemble_destroy": (None, [_P]),
    "ensemble_last_error": (C.c_char_p, [_P]),
    "ensemble_size": (_I, [_P]),
    "ensemble_chamber": (_P, [_P, _I]),
    "ensemble_step": (_I, [_P, _I]),
    "ensemble_sync": (_I, [_P]),
    "ensemble_autotune": (_I, [_P, _I]),
    "ensemble_tuning": (_I, [_P, C.c_char_p, _I]),
    "ensemble_profile_enable": (_I, [_P, _I]),
    "ensemble_profile_read": (_I, [_P, _P, _I, _P, _P, _I, C.POINTER(_I)]),
    "ensemble_profile_reset": (_I, [_P]),
    "ensemble_stream": (_P, [_P]),
    "surrogate_desc_default": (None, [_P]),
    "surrogate_create": (_I, [_P, C.POINTER(_P)]),
    "surrogate_destroy": (None, [_P]),
    "surrogate_last_error": (C.c_char_p, [_P]),
    "surrogate_train": (_I, [_P, _I, _I, C.c_float, C.c_uint64, C.POINTER(_D)]),
    "surrogate_sample": (_I, [_P, C.c_uint64]),
    "surrogate_set_state": (_I, [_P, _P, _P, _P, _P, C.c_size_t]),
    "surrogate_get_state": (_I, [_P, _P, _P, _P, _P, C.c_size_t]),
    "surrogate_advance": (_I, [_P, _I, C.POINTER(_I)]),
    "surrogate_counters_get": (_I, [_P, _P]),
    "surrogate_weights": (_I, [_P, _I, _P, C.c_size_t, _I]),
    "surrogate_eval": (_I, [_P, _P, C.c_size_t, _P, _P, _P, C.c_size_t]),
    "surrogate_stream": (_P, [_P]),
    "scenario_parse": (_I, [C.c_char_p, C.c_size_t, C.POINTER(Scenario)]),
    "scenario_load": (_I, [C.c_char_p, C.POINTER(Scenario)]),
    "scenario_capacities": (None, [C.POINTER(Scenario), C.POINTER(_I), C.POINTER(_I)]),
    "scenario_instantiate": (_I, [C.POINTER(Scenario), C.POINTER(WorldDesc), C.POINTER(_P)]),
    "scenario_apply_drive": (_I, [_P, C.POINTER(Scenario)]),
    "scenario_run_assert": (_I, [C.POINTER(Scenario), _P, C.c_int64, C.c_char_p, C.POINTER(AcceptResult), _I,
                                 C.POINTER(AssertSummary)]),
    "accept_eval": (_I, [C.POINTER(AcceptCheck), _D]),
    "metric_name": (C.c_char_p, [_I]),
}


class Binding:
    """The ABI of one shared library whose symbols carry `prefix`."""

    def __init__(self, path: Path | str, prefix: str = "astro_", extra: dict | None = None):
        path = Path(path)
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built")
        self.path = path
        self.lib = C.CDLL(str(path))
        self.prefix = prefix
        sigs = dict(SIGNATURES)
        if extra:
            sigs.update(extra)
        for name, (res, args) in sigs.items():
            fn = getattr(self.lib, prefix + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, "_" + name if name.startswith("philox_") else name, fn)

    def philox_blocks(self, words: np.ndarray) -> np.ndarray:
        """Philox4x32-10 of each row (c0, c1, c2, c3, k0, k1) of `words`."""
        w = np.ascontiguousarray(words, np.uint32).reshape(-1, 6)
        out = np.zeros((len(w), 4), np.uint32)
        st = self._philox_blocks(_ptr(w), len(w), _ptr(out))
        if st != OK:
            raise AstroError(st, "philox_blocks")
        return out

    def philox_stream(self, seed: int, cid: int, tick: int, stage: int, n: int) -> np.ndarray:
        out = np.zeros(n, np.uint32)
        st = self._philox_stream(seed, cid, tick, stage, n, _ptr(out))
        if st != OK:
            raise AstroError(st, "philox_stream")
        return out

    def desc(self, **kw) -> WorldDesc:
        """A WorldDesc with the reference defaults, overridden by `kw`.

        `grid` takes {field_id: (nx, ny, nz)}."""
        d = WorldDesc()
        self.world_desc_default(C.byref(d))
        grids = kw.pop("grid", {})
        for k, v in kw.items():
            if not hasattr(d, k):
                raise KeyError(k)
            setattr(d, k, v)
        for f, dims in grids.items():
            d.grid[f].nx, d.grid[f].ny, d.grid[f].nz = dims
        return d


_PRODUCT: Binding | None = None


def load() -> Binding:
    """Binds libastro_hot.so. Raises if it is not built or no GPU is usable."""
    global _PRODUCT
    if _PRODUCT is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python astrophage-microscope_amd/build.py`")
        _PRODUCT = Binding(LIB_PATH, "astro_", PRODUCT_ONLY)
    return _PRODUCT


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class World:
    """One simulated chamber (src/sim/world.cuh:89-180) behind the C ABI."""

    def __init__(self, binding: Binding, desc: WorldDesc | None = None, _handle=None, **kw):
        self.b = binding
        if _handle is None:
            if desc is None:
                desc = binding.desc(**kw)
            h = C.c_void_p()
            self._check(binding.world_create(C.byref(desc), C.byref(h)), what="world_create")
            _handle = h
        self.h = _handle

    @property
    def desc(self) -> WorldDesc:
        """The world's current description (clock and light included)."""
        d = WorldDesc()
        self._check(self.b.world_desc_get(self.h, C.byref(d)))
        return d

    # -- ASPH snapshots (src/sim/snapshot.h:17-40) ---------------------------
    @classmethod
    def from_snapshot(cls, binding: Binding, data: bytes, desc: WorldDesc | None = None, **kw) -> "World":
        """snapshot_from_bytes: a new world holding the snapshot's state; the
        configuration the file does not carry (grids, motion, light) comes
        from `desc` or `kw` as for World()."""
        if desc is None:
            desc = binding.desc(**kw)
        buf = C.create_string_buffer(bytes(data), len(data))
        h = C.c_void_p()
        st = binding.snapshot_from_bytes(C.byref(desc), C.cast(buf, C.c_void_p), len(data), C.byref(h))
        if st != OK:
            raise AstroError(st, "snapshot_from_bytes")
        return cls(binding, _handle=h)

    @classmethod
    def load_snapshot(cls, binding: Binding, path, desc: WorldDesc | None = None, **kw) -> "World":
        """snapshot_load."""
        if desc is None:
            desc = binding.desc(**kw)
        h = C.c_void_p()
        st = binding.snapshot_load(C.byref(desc), str(path).encode(), C.byref(h))
        if st != OK:
            raise AstroError(st, f"snapshot_load {path}")
        return cls(binding, _handle=h)

    def snapshot(self, scenario_id: str | None = None) -> bytes:
        """snapshot_to_bytes."""
        sid = scenario_id.encode() if scenario_id is not None else None
        n = C.c_size_t()
        self._check(self.b.snapshot_to_bytes(self.h, sid, None, 0, C.byref(n)), "snapshot size")
        buf = C.create_string_buffer(n.value)
        self._check(self.b.snapshot_to_bytes(self.h, sid, C.cast(buf, C.c_void_p), n.value, C.byref(n)),
                    "snapshot")
        return buf.raw[:n.value]

    def save_snapshot(self, path, scenario_id: str | None = None):
        """snapshot_save."""
        sid = scenario_id.encode() if scenario_id is not None else None
        self._check(self.b.snapshot_save(self.h, str(path).encode(), sid), f"snapshot_save {path}")

    # -- plumbing ----------------------------------------------------------
    def _check(self, st: int, what: str = ""):
        if st != OK:
            detail = ""
            if getattr(self, "h", None):
                detail = (self.b.last_error(self.h) or b"").decode()
            raise AstroError(st, detail or what)

    def close(self):
        if getattr(self, "h", None):
            self.b.world_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- population --------------------------------------------------------
    def spawn(self, count: int, seed: int, placement: int = PLACE_UNIFORM, place_x: float = 0.0,
              place_y: float = 0.0, place_radius: float = 0.0, charge_dist: int = DIST_CONSTANT,
              charge_a: float = 0.0, charge_b: float = 0.0, awake: bool = False):
        p = SpawnParams(count=count, placement=placement, place_x=place_x, place_y=place_y,
                        place_radius=place_radius, charge_dist=charge_dist, awake=int(awake),
                        charge_a=charge_a, charge_b=charge_b)
        self._check(self.b.cells_spawn(self.h, C.byref(p), C.c_uint64(seed)), "spawn")

    def upload(self, cells: dict[str, np.ndarray]):
        n = len(cells["id"])
        arrays = {}
        soa = CellSoA()
        for name, dt in CELL_FIELDS:
            a = cells.get(name)
            if a is None:
                a = np.zeros(n, dt)
            a = np.ascontiguousarray(a, dtype=dt)
            if len(a) != n:
                raise ValueError(f"{name} has {len(a)} entries, expected {n}")
            arrays[name] = a
            setattr(soa, name, _ptr(a))
        self._check(self.b.cells_upload(self.h, C.byref(soa), n), "upload")

    def download(self) -> dict[str, np.ndarray]:
        n = self.counters().count
        out = {name: np.zeros(n, dt) for name, dt in CELL_FIELDS}
        soa = CellSoA()
        for name, _ in CELL_FIELDS:
            setattr(soa, name, _ptr(out[name]))
        got = C.c_int32()
        self._check(self.b.cells_download(self.h, C.byref(soa), n, C.byref(got)), "download")
        return out

    def counters(self) -> Counters:
        c = Counters()
        self._check(self.b.world_get_counters(self.h, C.byref(c)))
        return c

    def set_counters(self, c: Counters):
        self._check(self.b.world_set_counters(self.h, C.byref(c)))

    # -- fields -----------------------------------------------------------
    def field_dims(self, f: int):
        g = GridDims()
        dx, dy, dz = C.c_double(), C.c_double(), C.c_double()
        self._check(self.b.field_dims(self.h, f, C.byref(g), C.byref(dx), C.byref(dy), C.byref(dz)))
        return (g.nx, g.ny, g.nz), (dx.value, dy.value, dz.value)

    def field(self, f: int) -> np.ndarray:
        (nx, ny, nz), _ = self.field_dims(f)
        a = np.zeros(nx * ny * nz, np.float32)
        self._check(self.b.field_download(self.h, f, _ptr(a), a.size), "field_download")
        return a.reshape(nz, ny, nx)

    def set_field(self, f: int, v: np.ndarray):
        a = np.ascontiguousarray(v, np.float32).ravel()
        self._check(self.b.field_upload(self.h, f, _ptr(a), a.size), "field_upload")

    def deposit(self, f: int) -> np.ndarray:
        (nx, ny, nz), _ = self.field_dims(f)
        a = np.zeros(nx * ny * nz, np.int64)
        self._check(self.b.field_deposit_download(self.h, f, _ptr(a), a.size))
        return a.reshape(nz, ny, nx)

    # -- stepping ----------------------------------------------------------
    def step(self, n: int = 1):
        self._check(self.b.world_step(self.h, n), "step")

    def sync(self):
        self._check(self.b.world_sync(self.h), "sync")

    def set_clock(self, physics: float, biology: float):
        self._check(self.b.world_set_clock(self.h, physics, biology))

    # -- Taumoeba -------------------------------------------------------------
    def tau_spawn(self, count: int, seed: int):
        self._check(self.b.taumoeba_spawn(self.h, count, seed), "taumoeba_spawn")

    def tau_counters(self) -> tuple[int, int, int]:
        """(count, capacity, next_id)."""
        n, cap, nid = C.c_int32(), C.c_int32(), C.c_uint64()
        self._check(self.b.taumoeba_counters(self.h, C.byref(n), C.byref(cap), C.byref(nid)))
        return n.value, cap.value, nid.value

    def tau_download(self) -> dict[str, np.ndarray]:
        n = self.tau_counters()[0]
        arrays = {name: np.zeros(n, dt) for name, dt in TAU_FIELDS}
        soa = TauSoA(*[_ptr(arrays[name]) for name, _ in TAU_FIELDS])
        got = C.c_int32()
        self._check(self.b.taumoeba_download(self.h, C.byref(soa), n, C.byref(got)), "taumoeba_download")
        return arrays

    def tau_upload(self, arrays: dict[str, np.ndarray]):
        n = len(arrays["id"])
        conv = {name: np.ascontiguousarray(arrays[name], dt) for name, dt in TAU_FIELDS}
        soa = TauSoA(*[_ptr(conv[name]) for name, _ in TAU_FIELDS])
        self._check(self.b.taumoeba_upload(self.h, C.byref(soa), n), "taumoeba_upload")

    def set_deposit_path(self, path: int):
        """Thermostat deposit path (DEPOSIT_AUTO / ATOMIC / DIRECT); the fields
        are the same bit for bit on each."""
        self._check(self.b.world_set_deposit_path(self.h, path))

    def set_clock_preset(self, preset: int, physics: float = 1.0, biology: float = 1.0):
        self._check(self.b.world_set_clock_preset(self.h, preset, physics, biology))

    def set_light(self, dir_x: float, dir_y: float, irradiance: float, ambient: float, enabled: bool):
        self._check(self.b.world_set_light(self.h, dir_x, dir_y, irradiance, ambient, 1 if enabled else 0))

    def brush(self, kind: int, x: float, y: float, radius: float, strength: float):
        self._check(self.b.world_apply_brush(self.h, kind, x, y, radius, strength), "brush")

    def set_flash(self, armed: bool):
        self._check(self.b.world_set_flash(self.h, 1 if armed else 0))

    def flash_audit(self) -> tuple[float, float]:
        """(|impulse| [N s], discharged energy [J])."""
        i, e = C.c_double(), C.c_double()
        self._check(self.b.world_flash_audit(self.h, C.byref(i), C.byref(e)))
        return i.value, e.value

    def fill(self, f: int, value: float):
        self._check(self.b.field_fill(self.h, f, value), "fill")

    def time(self) -> tuple[int, float, float, float]:
        """(tick, sim_time_s, physics_rate, biology_rate) without a device wait."""
        t, s, p, b = C.c_uint64(), C.c_double(), C.c_double(), C.c_double()
        self._check(self.b.world_time(self.h, C.byref(t), C.byref(s), C.byref(p), C.byref(b)))
        return t.value, s.value, p.value, b.value

    @property
    def dt(self) -> float:
        return 1.0e-3 * self.desc.physics_rate

    def stats(self) -> dict:
        s = Stats()
        self._check(self.b.world_stats(self.h, C.byref(s)), "stats")
        return {name: getattr(s, name) for name, _ in Stats._fields_ if not name.startswith("_")}

    def state_hash(self) -> int:
        v = C.c_uint64()
        self._check(self.b.world_state_hash(self.h, C.byref(v)), "state_hash")
        return v.value

    # -- stages --------------------------------------------------------------
    def hash_dims(self):
        nx, ny, nz, cell = C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
        self._check(self.b.hash_dims(self.h, C.byref(nx), C.byref(ny), C.byref(nz), C.byref(cell)))
        return nx.value, ny.value, nz.value, cell.value

    def hash_build(self):
        self._check(self.b.stage_hash_build(self.h), "hash_build")

    def hash_download(self):
        n = self.counters().count
        nx, ny, nz, _ = self.hash_dims()
        nb = nx * ny * nz
        vals = np.zeros(n, np.uint32)
        bs = np.zeros(nb, np.int32)
        be = np.zeros(nb, np.int32)
        self._check(self.b.hash_download(self.h, _ptr(vals), n, _ptr(bs), _ptr(be), nb))
        return vals, bs, be

    def stage(self, name: str, *args):
        fn = getattr(self.b, "stage_" + name)
        self._check(fn(self.h, *args), name)

    def forces(self):
        n = self.counters().count
        fx, fy, fz = (np.zeros(n) for _ in range(3))
        self._check(self.b.forces_download(self.h, _ptr(fx), _ptr(fy), _ptr(fz), n))
        return fx, fy, fz

    # -- measurement (product library only) --------------------------------
    def profile(self, on: bool):
        self._check(self.b.profile_enable(self.h, int(on)))

    def profile_reset(self):
        self._check(self.b.profile_reset(self.h))

    def profile_read(self, max_kernels: int = 64) -> dict[str, tuple[int, float]]:
        stride = 64
        names = C.create_string_buffer(stride * max_kernels)
        launches = np.zeros(max_kernels, np.int64)
        ms = np.zeros(max_kernels, np.float64)
        n = C.c_int32()
        self._check(self.b.profile_read(self.h, names, stride, _ptr(launches), _ptr(ms), max_kernels, C.byref(n)))
        out = {}
        for k in range(n.value):
            name = names.raw[k * stride:(k + 1) * stride].split(b"\0", 1)[0].decode()
            out[name] = (int(launches[k]), float(ms[k]))
        return out


def _read_profile(fn, h, max_kernels: int) -> dict[str, tuple[int, float]]:
    stride = 64
    names = C.create_string_buffer(stride * max_kernels)
    launches = np.zeros(max_kernels, np.int64)
    ms = np.zeros(max_kernels, np.float64)
    n = C.c_int32()
    st = fn(h, names, stride, _ptr(launches), _ptr(ms), max_kernels, C.byref(n))
    if st != OK:
        raise AstroError(st, "profile_read")
    out = {}
    for k in range(n.value):
        name = names.raw[k * stride:(k + 1) * stride].split(b"\0", 1)[0].decode()
        out[name] = (int(launches[k]), float(ms[k]))
    return out


class Ensemble:
    """Independent chambers stepped together: one batched launch per stage
    over every chamber, one graph launch per step.

    The reference has no ensemble; each chamber's results are those of
    World.step on that world alone. The worlds stay owned by the caller and
    must outlive the ensemble."""

    def __init__(self, binding: Binding, worlds: list[World]):
        self.b = binding
        self.worlds = list(worlds)
        arr = (C.c_void_p * len(self.worlds))(*[w.h.value for w in self.worlds])
        h = C.c_void_p()
        st = binding.ensemble_create(arr, len(self.worlds), C.byref(h))
        if st != OK:
            raise AstroError(st, "ensemble_create: chambers must agree on precision, clock, stages, fields "
                                 "and substep counts")
        self.h = h

    def _check(self, st: int, what: str = ""):
        if st != OK:
            detail = (self.b.ensemble_last_error(self.h) or b"").decode() if getattr(self, "h", None) else ""
            raise AstroError(st, detail or what)

    def close(self):
        if getattr(self, "h", None):
            self.b.ensemble_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self) -> int:
        return int(self.b.ensemble_size(self.h))

    def step(self, n: int = 1):
        self._check(self.b.ensemble_step(self.h, n), "ensemble_step")

    def sync(self):
        self._check(self.b.ensemble_sync(self.h), "ensemble_sync")

    def autotune(self, ticks_per_trial: int = 2) -> dict:
        """Times candidate launch shapes per stage on real ticks and keeps the
        fastest; the chambers advance by eleven trials of `ticks_per_trial`."""
        self._check(self.b.ensemble_autotune(self.h, ticks_per_trial), "ensemble_autotune")
        return self.tuning()

    def tuning(self) -> dict:
        buf = C.create_string_buffer(8192)
        self._check(self.b.ensemble_tuning(self.h, buf, len(buf)), "ensemble_tuning")
        return json.loads(buf.value.decode())

    def profile(self, on: bool):
        self._check(self.b.ensemble_profile_enable(self.h, int(on)))

    def profile_reset(self):
        self._check(self.b.ensemble_profile_reset(self.h))

    def profile_read(self, max_kernels: int = 64) -> dict[str, tuple[int, float]]:
        return _read_profile(self.b.ensemble_profile_read, self.h, max_kernels)


class SurrogateDesc(C.Structure):
    _fields_ = [("abi_version", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("d_temp", C.c_double), ("d_nutrient", C.c_double), ("dt_sub", C.c_double),
                ("steps_per_call", C.c_int32), ("cells", C.c_int32), ("uptake", C.c_double),
                ("temp_ref", C.c_double), ("nutrient_ref", C.c_double), ("seed", C.c_uint64),
                ("leak", C.c_double), ("judge_every", C.c_int32), ("judge_tol", C.c_double), ("budget_tol", C.c_double),
                ("cooldown", C.c_int32)]


class SurrogateCounters(C.Structure):
    _fields_ = [("calls", C.c_int64), ("fallbacks", C.c_int64), ("judged", C.c_int64),
                ("last_judge_err", C.c_double), ("last_budget_residual", C.c_double),
                ("out_scale_temp", C.c_double), ("out_scale_nutrient", C.c_double)]


SURROGATE_PLAIN, SURROGATE_NET, SURROGATE_GATED = 0, 1, 2
SURROGATE_CIN = [8, 32, 32, 32]


class Surrogate:
    """The learned field surrogate (BF16 MFMA 3D-conv operator) behind the C
    ABI; the reference has none. Fields are float32 arrays [nz][ny][nx]."""

    def __init__(self, binding: Binding, **kw):
        self.b = binding
        d = SurrogateDesc()
        binding.surrogate_desc_default(C.byref(d))
        for k, v in kw.items():
            if not hasattr(d, k):
                raise KeyError(k)
            setattr(d, k, v)
        h = C.c_void_p()
        st = binding.surrogate_create(C.byref(d), C.byref(h))
        if st != OK:
            raise AstroError(st, "surrogate_create")
        self.h = h
        self.desc = d
        self.shape = (d.nz, d.ny, d.nx)
        self.n = d.nx * d.ny * d.nz

    def _check(self, st: int, what: str = ""):
        if st != OK:
            detail = (self.b.surrogate_last_error(self.h) or b"").decode() if getattr(self, "h", None) else ""
            raise AstroError(st, detail or what)

    def close(self):
        if getattr(self, "h", None):
            self.b.surrogate_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> dict:
        d = self.desc
        return {k: getattr(d, k) for k in ("dx", "dy", "dz", "dt_sub", "d_temp", "d_nutrient", "uptake",
                                           "steps_per_call")}

    def train(self, trajectories: int, epochs: int = 1, lr: float = 2e-3, seed: int = 1) -> float:
        loss = C.c_double()
        self._check(self.b.surrogate_train(self.h, trajectories, epochs, lr, seed, C.byref(loss)), "train")
        return loss.value

    def sample(self, seed: int):
        self._check(self.b.surrogate_sample(self.h, seed), "sample")

    def set_state(self, T=None, Cn=None, S=None, occ=None):
        arr = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (T, Cn, S, occ)]
        self._check(self.b.surrogate_set_state(self.h, *[_ptr(a) for a in arr], self.n), "set_state")

    def state(self):
        out = [np.empty(self.shape, np.float32) for _ in range(4)]
        self._check(self.b.surrogate_get_state(self.h, *[_ptr(a) for a in out], self.n), "get_state")
        return tuple(out)

    def advance(self, mode: int = SURROGATE_GATED) -> bool:
        used = C.c_int32()
        self._check(self.b.surrogate_advance(self.h, mode, C.byref(used)), "advance")
        return bool(used.value)

    def counters(self) -> dict:
        c = SurrogateCounters()
        self._check(self.b.surrogate_counters_get(self.h, C.byref(c)))
        return {name: getattr(c, name) for name, _ in SurrogateCounters._fields_}

    def weights(self, layer: int) -> np.ndarray:
        w = np.empty(27 * 32 * SURROGATE_CIN[layer] + 32, np.float32)
        self._check(self.b.surrogate_weights(self.h, layer, _ptr(w), w.size, 0))
        return w

    def set_weights(self, layer: int, w: np.ndarray):
        w = np.ascontiguousarray(w, np.float32)
        self._check(self.b.surrogate_weights(self.h, layer, _ptr(w), w.size, 1))

    def eval(self, x: np.ndarray, target: np.ndarray | None = None, grads: bool = False):
        """Network on normalized inputs x [..][8]: predictions [..][2], and
        with a target the per-layer weight gradients of the mean squared
        error."""
        x = np.ascontiguousarray(x, np.float32)
        pred = np.empty(x.shape[:-1] + (2,), np.float32)
        tgt = None if target is None else np.ascontiguousarray(target, np.float32)
        sizes = [27 * 32 * c + 32 for c in SURROGATE_CIN]
        g = np.empty(sum(sizes), np.float32) if grads else None
        self._check(self.b.surrogate_eval(self.h, _ptr(x), self.n, _ptr(tgt), _ptr(pred), _ptr(g),
                                          0 if g is None else g.size), "eval")
        if not grads:
            return pred
        out, off = [], 0
        for sz in sizes:
            out.append(g[off:off + sz])
            off += sz
        return pred, out


# ---- scenarios ----------------------------------------------------------------
def scenario_parse(binding: Binding, text: str | bytes) -> Scenario:
    """scenario_parse: raises AstroError on a malformed scenario."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    sc = Scenario()
    st = binding.scenario_parse(data, len(data), C.byref(sc))
    if st != OK:
        raise AstroError(st, "scenario_parse")
    return sc


def scenario_load(binding: Binding, path) -> Scenario:
    sc = Scenario()
    st = binding.scenario_load(str(path).encode(), C.byref(sc))
    if st != OK:
        raise AstroError(st, f"scenario_load {path}")
    return sc


def scenario_world(binding: Binding, sc: Scenario, desc: WorldDesc | None = None, **kw) -> World:
    """scenario_instantiate: a new world; grids, precision and the hash cell
    come from `desc` or `kw` (as for World())."""
    if desc is None:
        desc = binding.desc(**kw)
    h = C.c_void_p()
    st = binding.scenario_instantiate(C.byref(sc), C.byref(desc), C.byref(h))
    if st != OK:
        raise AstroError(st, "scenario_instantiate")
    return World(binding, _handle=h)


def scenario_drive(binding: Binding, w: World, sc: Scenario):
    """scenario_apply_drive for the world's current tick."""
    w._check(binding.scenario_apply_drive(w.h, C.byref(sc)), "scenario_apply_drive")


def scenario_assert(binding: Binding, sc: Scenario, w: World, ticks: int = 0, csv_path=None) -> dict:
    """run_scenario_assert: {'all_pass', 'ticks', 'sim_time_s', 'samples',
    'checks': [{'metric', 'op', 'measured', 'value', 'tol', 'pass'}]}."""
    res = (AcceptResult * 16)()
    summ = AssertSummary()
    csv = str(csv_path).encode() if csv_path is not None else None
    w._check(binding.scenario_run_assert(C.byref(sc), w.h, ticks, csv, res, 16, C.byref(summ)), "run_assert")
    checks = []
    for i in range(min(summ.checks, 16)):
        r = res[i]
        checks.append(dict(metric=METRICS[r.metric], op=OPS[r.op], measured=r.measured, value=r.value, tol=r.tol,
                           tol_absolute=bool(r.tol_absolute), after_s=r.after_s, passed=bool(r.pass_)))
    return dict(all_pass=bool(summ.all_pass), ticks=summ.ticks, sim_time_s=summ.sim_time_s,
                samples=summ.samples, checks=checks)
