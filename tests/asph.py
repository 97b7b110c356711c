"""The ASPH v1 snapshot file, read independently of the library for tests.

Layout (contracts/snapshot_v1.h:14-57): a 256-byte little-endian header, then
`override_count` ParamOverride records ({char key[48]; double value}), the
cell SoA [0, count) array by array in contract order
(contracts/cell_store_v1.h:42-78), the Taumoeba arrays
(src/sim/snapshot.cu:59-67), and the
T, CO2, N2 values as float32. state_hash is FNV-1a 64 over everything after
the header. The header format below is checked field by field against the
reference's own struct in tests/test_snapshot_format.py.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = 0x48505341   # 'ASPH'
VERSION = 1
# The contract's FNV-1a 64 starts from 1469598103934665603
# (contracts/snapshot_v1.h:59), not the published offset basis
# 14695981039346656037 (one digit longer); the files carry the former.
FNV_OFFSET = 1469598103934665603
FNV_PUBLISHED_OFFSET = 14695981039346656037
FNV_PRIME = 0x100000001B3

# (name, struct code) in declaration order; "x" entries are padding.
HEADER = [
    ("magic", "I"), ("version", "I"), ("seed", "Q"), ("tick", "Q"), ("sim_time_s", "d"),
    ("cell_count", "i"), ("cell_capacity", "i"), ("taumoeba_count", "i"), ("taumoeba_capacity", "i"),
    ("n_temp", "i"), ("n_co2", "i"), ("n_n2", "i"), (None, "4x"),
    ("chamber_w", "d"), ("chamber_h", "d"), ("chamber_d", "d"), ("physics_rate", "d"), ("biology_rate", "d"),
    ("next_cell_id", "Q"), ("override_count", "i"), ("non_canon_run", "B"), (None, "3x"),
    ("scenario_id", "64s"), ("build_describe", "64s"), ("state_hash", "Q"),
]
HEADER_FMT = "<" + "".join(code for _, code in HEADER)
HEADER_SIZE = struct.calcsize(HEADER_FMT)
OVERRIDE_FMT = "<48sd"
OVERRIDE_SIZE = struct.calcsize(OVERRIDE_FMT)


def header_offsets() -> dict[str, int]:
    """Byte offset of every named header field."""
    out, fmt = {}, "<"
    for name, code in HEADER:
        if name is not None:
            out[name] = struct.calcsize(fmt)
        fmt += code
    return out


def fnv1a64(data: bytes, h: int = FNV_OFFSET) -> int:
    for b in data:
        h = ((h ^ b) * FNV_PRIME) & 0xFFFFFFFFFFFFFFFF
    return h


def field_values(n: int) -> int:
    """Values stored for a field whose header size is n: n*n for the
    reference's square 2D grid, -n for a 3D grid, none when 0."""
    return n * n if n > 0 else -n


def parse(data: bytes, cell_fields, tau_fields=None) -> dict:
    """Header, overrides, cell arrays, Taumoeba arrays and field values of a
    snapshot. `cell_fields` and `tau_fields` are [(name, numpy dtype)] in
    the file's order."""
    if len(data) < HEADER_SIZE:
        raise ValueError("shorter than a header")
    vals = struct.unpack_from(HEADER_FMT, data, 0)
    names = [n for n, _ in HEADER if n is not None]
    h = dict(zip(names, vals))
    for k in ("scenario_id", "build_describe"):
        h[k] = h[k].split(b"\0", 1)[0].decode()
    off = HEADER_SIZE
    overrides = {}
    for _ in range(h["override_count"]):
        key, value = struct.unpack_from(OVERRIDE_FMT, data, off)
        overrides[key.split(b"\0", 1)[0].decode()] = value
        off += OVERRIDE_SIZE
    n = h["cell_count"]
    cells = {}
    for name, dt in cell_fields:
        a = np.frombuffer(data, dtype=np.dtype(dt).newbyteorder("<"), count=n, offset=off)
        cells[name] = a.copy()
        off += a.nbytes
    tau = {}
    if h["taumoeba_count"] != 0:
        if tau_fields is None:
            raise ValueError("the snapshot holds Taumoeba: pass their fields")
        for name, dt in tau_fields:
            a = np.frombuffer(data, dtype=np.dtype(dt).newbyteorder("<"), count=h["taumoeba_count"], offset=off)
            tau[name] = a.copy()
            off += a.nbytes
    fields = []
    for key in ("n_temp", "n_co2", "n_n2"):
        m = field_values(h[key])
        fields.append(np.frombuffer(data, dtype="<f4", count=m, offset=off).copy())
        off += 4 * m
    if off != len(data):
        raise ValueError(f"{len(data) - off} trailing bytes")
    return dict(header=h, overrides=overrides, cells=cells, tau=tau, fields=fields)


def body_of(overrides: dict[str, float], cells: dict, cell_fields, fields) -> bytes:
    """The body a world with this state serializes to."""
    out = bytearray()
    for key, value in overrides.items():
        out += struct.pack(OVERRIDE_FMT, key.encode(), value)
    for name, dt in cell_fields:
        out += np.ascontiguousarray(cells[name], np.dtype(dt).newbyteorder("<")).tobytes()
    for f in fields:
        out += np.ascontiguousarray(f, "<f4").tobytes()
    return bytes(out)
