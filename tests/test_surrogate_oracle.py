"""The surrogate's CPU oracle pinned by what the mathematics fixes: the plain
dynamics conserve heat up to the source, pure uptake has a closed form, the
FTCS stencil spreads a Gaussian's variance by exactly 2 c dx^2 per substep
and per axis away from walls, a constant field is a fixed point; the network
gradients agree with central finite differences of the loss."""
import sys

import numpy as np

from conftest import ROOT

sys.path.insert(0, str(ROOT / "oracle"))
import surrogate_oracle as so  # noqa: E402

D = dict(dx=4.3e-6, dy=4.3e-6, dz=1.56e-6, dt_sub=6.25e-6, d_temp=1.4325e-7, d_nutrient=1.92e-9, uptake=200.0,
         steps_per_call=16)


def test_heat_conserved_up_to_source():
    rng = np.random.default_rng(1)
    T = 369.0 + rng.normal(0, 2, (12, 10, 16))
    S = rng.uniform(0, 300, T.shape)
    C = np.full(T.shape, 1e-3)
    occ = np.zeros(T.shape)
    T1, _ = so.plain_advance(T, C, S, occ, D)
    want = T.sum() + D["steps_per_call"] * D["dt_sub"] * S.sum()
    assert abs(T1.sum() - want) <= 1e-12 * abs(want)


def test_pure_uptake_closed_form():
    rng = np.random.default_rng(2)
    C = rng.uniform(1e-4, 2e-3, (6, 7, 8))
    occ = rng.integers(0, 4, C.shape).astype(float)
    d = dict(D, d_nutrient=0.0)
    _, C1 = so.plain_advance(np.zeros(C.shape), C, np.zeros(C.shape), occ, d)
    want = C * (1.0 - d["dt_sub"] * d["uptake"] * occ) ** d["steps_per_call"]
    assert np.allclose(C1, want, rtol=1e-13, atol=0)


def test_gaussian_variance_grows_by_2c_per_substep():
    n = 64
    z, y, x = np.meshgrid(np.arange(16), np.arange(n), np.arange(n), indexing="ij")
    u = np.exp(-((x - 31.5) ** 2 + (y - 31.5) ** 2) / (2 * 3.0 ** 2)) * (1 + 0 * z)
    cx, cy = 0.11, 0.07

    def var(f, coord):
        m = f.sum()
        mu = (f * coord).sum() / m
        return (f * (coord - mu) ** 2).sum() / m

    v0x, v0y = var(u, x), var(u, y)
    for _ in range(10):
        u = so.ftcs_substep(u, cx, cy, 0.2)
    assert abs(var(u, x) - (v0x + 10 * 2 * cx)) < 1e-9
    assert abs(var(u, y) - (v0y + 10 * 2 * cy)) < 1e-9


def test_constant_field_is_fixed_point():
    u = np.full((5, 6, 7), 370.25)
    assert np.array_equal(so.ftcs_substep(u, 0.1, 0.1, 0.25), u)


def test_network_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    shape = (3, 3, 4)
    x = np.zeros(shape + (8,))
    x[..., :5] = rng.normal(0, 1, shape + (5,))
    tgt = rng.normal(0, 1, shape + (2,))
    cin = [8, 32, 32, 32]
    layers = []
    for li, c in enumerate(cin):
        v = rng.normal(0, 0.2, 27 * 32 * c + 32)
        if li == 3:   # only two outputs exist
            w = v[:27 * 32 * c].reshape(27, 32, c)
            w[:, 2:, :] = 0
            v[27 * 32 * c + 2:] = 0
        layers.append(v)
    params = so.unpack(layers, cin)
    loss, _, grads = so.loss_and_grads(x, tgt, params, rounding=False)
    h = 1e-6
    for li, idx in [(0, (4, 3, 2)), (1, (13, 7, 30)), (2, (0, 31, 5)), (3, (26, 1, 17))]:
        w, b = params[li]
        wp = [(ww.copy(), bb.copy()) for ww, bb in params]
        wm = [(ww.copy(), bb.copy()) for ww, bb in params]
        wp[li][0][idx] += h
        wm[li][0][idx] -= h
        fd = (so.loss_and_grads(x, tgt, wp, rounding=False)[0] - so.loss_and_grads(x, tgt, wm, rounding=False)[0]) / (2 * h)
        assert abs(fd - grads[li][0][idx]) <= 1e-6 * max(1.0, abs(fd)), (li, fd, grads[li][0][idx])
    bp = [(ww.copy(), bb.copy()) for ww, bb in params]
    bm = [(ww.copy(), bb.copy()) for ww, bb in params]
    bp[2][1][9] += h
    bm[2][1][9] -= h
    fd = (so.loss_and_grads(x, tgt, bp, rounding=False)[0] - so.loss_and_grads(x, tgt, bm, rounding=False)[0]) / (2 * h)
    assert abs(fd - grads[2][1][9]) <= 1e-6 * max(1.0, abs(fd))
    assert loss > 0


def test_bf16_rounding_ties_to_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-3], np.float32)
    r = so.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even (1.0)
    assert r[2] == 1.0 + 2 ** -6            # tie -> even (1 + 4 * 2^-8)
    assert r[3] == 1.0
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8
