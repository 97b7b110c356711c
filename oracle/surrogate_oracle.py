"""CPU oracle of the field surrogate's path (test infrastructure only: used by
tests/ and by bench.py's CPU-reference leg as the checker, never by the
product).

Two parts, both FP64 numpy:
  * plain_advance: the dynamics one surrogate call replaces. Per substep both
    fields take one 7-point FTCS substep with Neumann walls (ghost = the
    cell's own value), the 3D form of the reference's diffuse_kernel
    (src/fields/diffuse.cu:19-44, stencil u' = u + c * (sum of neighbours -
    2u per axis)); then T += dt * S and C -= dt * uptake * occupancy * C.
    Pinned in tests/test_surrogate_oracle.py by what the mathematics fixes:
    exact heat conservation plus the source, the closed form of pure uptake,
    Gaussian spreading sigma^2 = sigma0^2 + 2 D t, and a constant field being
    a fixed point.
  * the network (forward and gradients): the 4-layer 3x3x3 convolution with
    zero padding and leaky ReLU (slope `leak`), with BF16 rounding at the same points as the
    device (inputs, weights, hidden activations), accumulating in FP64.
    Pinned against central finite differences of its own loss.
Arrays are [z][y][x] (fields) and [z][y][x][channel] (activations).
"""
from __future__ import annotations

import numpy as np

LEAK = 0.1
TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


# --------------------------------------------------------------------------
# plain dynamics
# --------------------------------------------------------------------------
def ftcs_substep(u: np.ndarray, cx: float, cy: float, cz: float) -> np.ndarray:
    p = np.pad(u, 1, mode="edge")   # Neumann ghost = the boundary cell's value
    c = p[1:-1, 1:-1, 1:-1]
    lx = p[1:-1, 1:-1, :-2] + p[1:-1, 1:-1, 2:] - 2.0 * c
    ly = p[1:-1, :-2, 1:-1] + p[1:-1, 2:, 1:-1] - 2.0 * c
    lz = p[:-2, 1:-1, 1:-1] + p[2:, 1:-1, 1:-1] - 2.0 * c
    return c + (cx * lx + cy * ly + cz * lz)


def coefficients(d: dict, D: float) -> tuple[float, float, float]:
    return (D * d["dt_sub"] / d["dx"] ** 2, D * d["dt_sub"] / d["dy"] ** 2, D * d["dt_sub"] / d["dz"] ** 2)


def plain_advance(T, C, S, occ, d: dict, steps: int | None = None):
    """One call of the plain path in FP64; d holds dx, dy, dz, dt_sub,
    d_temp, d_nutrient, uptake, steps_per_call."""
    T = np.asarray(T, np.float64).copy()
    C = np.asarray(C, np.float64).copy()
    S = np.asarray(S, np.float64)
    occ = np.asarray(occ, np.float64)
    ct = coefficients(d, d["d_temp"])
    cc = coefficients(d, d["d_nutrient"])
    dt = d["dt_sub"]
    for _ in range(d["steps_per_call"] if steps is None else steps):
        T = ftcs_substep(T, *ct)
        C = ftcs_substep(C, *cc)
        T = T + dt * S
        C = C - dt * d["uptake"] * occ * C
    return T, C


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


# --------------------------------------------------------------------------
# the network
# --------------------------------------------------------------------------
def bf16(x) -> np.ndarray:
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def conv(a: np.ndarray, w: np.ndarray) -> np.ndarray:
    """a [z][y][x][cin], w [27][cout][cin] -> [z][y][x][cout], zero padding."""
    nz, ny, nx, _ = a.shape
    p = np.pad(a, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((nz, ny, nx, w.shape[1]))
    for t, (dz, dy, dx) in enumerate(TAPS):
        sl = p[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx, :]
        out += sl @ w[t].T
    return out


def conv_transpose(g: np.ndarray, w: np.ndarray) -> np.ndarray:
    """The input gradient of conv: g [z][y][x][cout] -> [z][y][x][cin]."""
    nz, ny, nx, _ = g.shape
    p = np.pad(g, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((nz, ny, nx, w.shape[2]))
    for t, (dz, dy, dx) in enumerate(TAPS):
        sl = p[1 - dz:1 - dz + nz, 1 - dy:1 - dy + ny, 1 - dx:1 - dx + nx, :]
        out += sl @ w[t]
    return out


def conv_wgrad(a: np.ndarray, g: np.ndarray) -> np.ndarray:
    """dW[t][cout][cin] = sum_v g[v][cout] a[v + off_t][cin]."""
    nz, ny, nx, cin = a.shape
    p = np.pad(a, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((27, g.shape[3], cin))
    gf = g.reshape(-1, g.shape[3])
    for t, (dz, dy, dx) in enumerate(TAPS):
        sl = p[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx, :].reshape(-1, cin)
        out[t] = gf.T @ sl
    return out


def unpack(layers: list[np.ndarray], cinp: list[int]):
    """Flat per-layer vectors ([27][32][cin] + [32] bias) -> (W, b) pairs."""
    out = []
    for v, c in zip(layers, cinp):
        w = np.asarray(v[:27 * 32 * c], np.float64).reshape(27, 32, c)
        b = np.asarray(v[27 * 32 * c:], np.float64)
        out.append((w, b))
    return out


def forward(x: np.ndarray, params, nout: int = 2, rounding: bool = True, leak: float = LEAK):
    """x [z][y][x][8] -> (pred [..][nout], activations for backward). With
    `rounding`, rounds as the device does: inputs, weights and hidden
    activations to BF16."""
    r = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    acts = [r(x)]
    zs = []
    a = acts[0]
    for li, (w, b) in enumerate(params):
        z = conv(a, r(w)) + b
        zs.append(z)
        if li < len(params) - 1:
            a = r(np.where(z > 0, z, leak * z))
            acts.append(a)
    return zs[-1][..., :nout], acts, zs


def loss_and_grads(x, target, params, nout: int = 2, rounding: bool = True, leak: float = LEAK):
    """Mean squared error over voxels and outputs, and its gradient with
    respect to every layer's (master) weights and biases (BF16 weights act as
    if their rounding were the identity, as in training)."""
    r = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    pred, acts, zs = forward(x, params, nout, rounding, leak)
    n = pred.size
    loss = float(np.mean((pred - target) ** 2))
    g = np.zeros(zs[-1].shape)
    g[..., :nout] = 2.0 * (pred - target) / n
    grads = [None] * len(params)
    for li in range(len(params) - 1, -1, -1):
        w, _ = params[li]
        dw = conv_wgrad(acts[li], g)
        db = g.reshape(-1, g.shape[-1]).sum(axis=0)
        grads[li] = (dw, db)
        if li > 0:
            da = conv_transpose(g, r(w))
            g = da * np.where(acts[li] > 0, 1.0, leak)
    return loss, pred, grads
