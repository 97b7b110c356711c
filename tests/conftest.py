"""Shared fixtures.

Three libraries are in play:
  * the HIP product, astrophage-microscope_amd/lib/libastro_hot.so;
  * the CPU oracle, oracle/liboracle.so -- a restatement of the reference's
    step used only as the checker;
  * oracle/_ref/libastro_ref.so -- the reference's own host-callable physics
    headers compiled from source, present only where the reference checkout is.
GPU tests are marked `gpu`; everything else runs on the CPU.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "astrophage-microscope_amd"
ORACLE = ROOT / "oracle" / "liboracle.so"
REF = ROOT / "oracle" / "_ref" / "libastro_ref.so"
sys.path.insert(0, str(PKG))
sys.path.insert(0, str(ROOT))

import astro_hot as ah  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _ensure_oracle():
    if not ORACLE.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle"), "oracle"], check=True,
                       stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def oracle() -> ah.Binding:
    _ensure_oracle()
    return ah.Binding(ORACLE, "oracle_")


@pytest.fixture(scope="session")
def ref():
    if not REF.exists():
        pytest.skip("reference build (oracle/_ref) not present")
    import ctypes as C
    return C.CDLL(str(REF))


@pytest.fixture(scope="session")
def gpu() -> ah.Binding:
    if not ah.LIB_PATH.exists():
        sys.path.insert(0, str(PKG))
        import build as _b  # noqa: F401
        _b.build()
    b = ah.load()
    return b


def oracle_cells(oracle_binding, n: int, spawn_seed: int, spawn: dict | None = None, **desc_kw) -> dict:
    """Cells spawned by the oracle (reference spawn order), as host arrays."""
    with ah.World(oracle_binding, **desc_kw) as w:
        w.spawn(n, spawn_seed, **(spawn or {}))
        return w.download()


def compare_cells(a: dict, b: dict, rtol: float = 0.0, atol: dict | None = None,
                  exact=("id", "flags", "death_cause", "rng_state")) -> dict:
    """Max abs difference per array; integer arrays must match exactly."""
    atol = atol or {}
    out = {}
    for name, _ in ah.CELL_FIELDS:
        x, y = a[name], b[name]
        assert x.shape == y.shape, (name, x.shape, y.shape)
        if name in exact:
            bad = np.nonzero(x != y)[0]
            assert bad.size == 0, f"{name} differs at {bad[:10]} ({x[bad[:5]]} vs {y[bad[:5]]})"
            out[name] = 0.0
        else:
            d = np.abs(x.astype(np.float64) - y.astype(np.float64))
            out[name] = float(d.max()) if d.size else 0.0
            if name in atol:
                assert out[name] <= atol[name], f"{name}: max |diff| {out[name]} > {atol[name]}"
    return out


os.environ.setdefault("HIP_LAUNCH_BLOCKING", "0")
