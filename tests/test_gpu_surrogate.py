"""The learned field surrogate on the GPU against its CPU oracle
(oracle/surrogate_oracle.py).

Bars:
  * the BF16 MFMA network: predictions within 5e-3 relative L2 of the FP64
    oracle that rounds at the same points (inputs, weights and hidden
    activations to BF16); weight gradients within 3e-2 relative L2 per layer
    (the device also rounds the backpropagated gradients to BF16);
  * predictions come from two sets of kernels: the inference kernels on the
    halo-padded layout (eval without a target, and every net call) and the
    training kernels (eval with a target); both are held to the oracle;
  * the plain path (one fused pass per substep: both fields' FTCS plus the
    sources), FP32 against the FP64 oracle: fields within 1e-6 relative L2,
    increments within 2e-3 (FP32 rounding of fields near 369 K), at a small
    shape and at the benchmarked 128^3 shape;
  * a gated call that the judge rejects returns exactly the plain result;
  * training brings the increment errors well below predicting no change.
"""
import sys

import numpy as np
import pytest

from conftest import ROOT, ah

sys.path.insert(0, str(ROOT / "oracle"))
import surrogate_oracle as so  # noqa: E402

pytestmark = pytest.mark.gpu

SMALL = dict(nx=32, ny=6, nz=5, cells=40, dx=550e-6 / 32, dy=550e-6 / 32, dz=200e-6 / 16, leak=0.1)


def random_layers(rng, scale=0.15):
    layers = []
    for li, c in enumerate(ah.SURROGATE_CIN):
        v = rng.normal(0, scale, 27 * 32 * c + 32).astype(np.float32)
        if li == 3:
            v[:27 * 32 * c].reshape(27, 32, c)[:, 2:, :] = 0
            v[27 * 32 * c + 2:] = 0
        layers.append(v)
    return layers


def random_inputs(rng, shape):
    x = rng.normal(0, 1, shape + (8,)).astype(np.float32)
    x[..., 4] = 1.0
    return x


@pytest.mark.parametrize("dims", [(32, 6, 5), (160, 5, 3)])
def test_network_forward_matches_oracle(gpu, dims):
    """(160, 5, 3): five 32-voxel x tiles, and row counts (6 and 5) that end
    in partial row groups of the inference kernels."""
    rng = np.random.default_rng(11)
    kw = dict(SMALL, nx=dims[0], ny=dims[1], nz=dims[2], dx=550e-6 / dims[0], dy=550e-6 / dims[0])
    with ah.Surrogate(gpu, **kw) as s:
        layers = random_layers(rng)
        for li, v in enumerate(layers):
            s.set_weights(li, v)
            assert np.array_equal(s.weights(li), v)
        x = random_inputs(rng, s.shape)
        pred = s.eval(x)
        want, _, _ = so.forward(x, so.unpack(layers, ah.SURROGATE_CIN), leak=0.1)
        err = so.rel_l2(pred, want)
        assert err < 5e-3, err


@pytest.mark.parametrize("dims", [(32, 6, 5), (160, 5, 3)])
def test_network_gradients_match_oracle(gpu, dims):
    rng = np.random.default_rng(12)
    kw = dict(SMALL, nx=dims[0], ny=dims[1], nz=dims[2], dx=550e-6 / dims[0], dy=550e-6 / dims[0])
    with ah.Surrogate(gpu, **kw) as s:
        layers = random_layers(rng)
        for li, v in enumerate(layers):
            s.set_weights(li, v)
        x = random_inputs(rng, s.shape)
        tgt = rng.normal(0, 1, s.shape + (2,)).astype(np.float32)
        pred, grads = s.eval(x, tgt, grads=True)
        _, want_pred, want = so.loss_and_grads(x, tgt, so.unpack(layers, ah.SURROGATE_CIN), leak=0.1)
        assert so.rel_l2(pred, want_pred) < 5e-3   # the training kernels' forward pass
        for li, (gw, gb) in enumerate(want):
            c = ah.SURROGATE_CIN[li]
            got_w = grads[li][:27 * 32 * c].reshape(27, 32, c)
            got_b = grads[li][27 * 32 * c:]
            nout = 2 if li == 3 else 32
            e_w = so.rel_l2(got_w[:, :nout, :], gw[:, :nout, :])
            e_b = so.rel_l2(got_b[:nout], gb[:nout])
            assert e_w < 3e-2 and e_b < 3e-2, (li, e_w, e_b)


def test_plain_path_matches_fp64_oracle(gpu):
    with ah.Surrogate(gpu, nx=64, ny=48, nz=24, cells=3000, dx=550e-6 / 64, dy=550e-6 / 64,
                      dz=200e-6 / 48) as s:
        s.sample(5)
        T0, C0, S, occ = s.state()
        assert occ.sum() == 3000
        s.advance(ah.SURROGATE_PLAIN)
        T1, C1, _, _ = s.state()
        Tw, Cw = so.plain_advance(T0, C0, S, occ, s.params())
        # FP32 fields near 369 K carry 3e-5 K per rounding; the increments
        # (tenths of a kelvin) inherit that: ~1e-3 relative at most
        assert so.rel_l2(T1, Tw) < 1e-6 and so.rel_l2(C1, Cw) < 1e-6
        assert so.rel_l2(T1.astype(np.float64) - T0, Tw - T0) < 2e-3
        assert so.rel_l2(C1.astype(np.float64) - C0, Cw - C0) < 2e-3


def test_plain_path_full_shape_matches_fp64_oracle(gpu):
    """The benchmarked shape: 128^3 over 550 x 550 x 200 um, 25,000 cells,
    16 substeps per call; two calls."""
    with ah.Surrogate(gpu) as s:
        assert s.shape == (128, 128, 128)
        s.sample(21)
        T0, C0, S, occ = s.state()
        assert occ.sum() == 25000
        s.advance(ah.SURROGATE_PLAIN)
        s.advance(ah.SURROGATE_PLAIN)
        T1, C1, _, _ = s.state()
        Tw, Cw = so.plain_advance(T0, C0, S, occ, s.params(), steps=2 * s.params()["steps_per_call"])
        assert so.rel_l2(T1, Tw) < 1e-6 and so.rel_l2(C1, Cw) < 1e-6
        assert so.rel_l2(T1.astype(np.float64) - T0, Tw - T0) < 2e-3
        assert so.rel_l2(C1.astype(np.float64) - C0, Cw - C0) < 2e-3


@pytest.mark.parametrize("gate", ["judge", "budget"])
def test_gated_call_rejected_is_the_plain_call(gpu, gate):
    """An untrained net is rejected either by the plain-path judge (budget
    gate off) or by the heat-budget gate (judge off)."""
    kw = dict(SMALL, nx=64, ny=16, nz=12, cells=500, cooldown=2)
    if gate == "judge":
        kw.update(judge_every=1, judge_tol=1e-9, budget_tol=1e30)
    else:
        kw.update(judge_every=0, budget_tol=1e-9)
    with ah.Surrogate(gpu, **kw) as s, ah.Surrogate(gpu, **kw) as p:
        s.sample(9)
        p.sample(9)
        assert not s.advance(ah.SURROGATE_GATED)   # untrained: rejected
        p.advance(ah.SURROGATE_PLAIN)
        for a, b in zip(s.state(), p.state()):
            assert np.array_equal(a, b)
        c = s.counters()
        assert c["fallbacks"] == 1
        if gate == "judge":
            assert c["judged"] == 1 and c["last_judge_err"] > 1e-9
        else:
            assert c["judged"] == 0 and c["last_budget_residual"] > 1e-9
        # the cooldown runs plain calls without asking the net
        assert not s.advance(ah.SURROGATE_GATED)
        p.advance(ah.SURROGATE_PLAIN)
        for a, b in zip(s.state(), p.state()):
            assert np.array_equal(a, b)


def test_training_lowers_increment_error(gpu):
    """Training on plain trajectories brings the net's T and nutrient
    increments well below the error of predicting no change (1.0)."""
    kw = dict(nx=64, ny=64, nz=32, cells=4000, dx=550e-6 / 64, dy=550e-6 / 64, dz=200e-6 / 64, seed=3)
    with ah.Surrogate(gpu, **kw) as s:
        def err():
            out = []
            for seed in (777, 778):
                s.sample(seed)
                T0, C0, S, occ = s.state()
                s.advance(ah.SURROGATE_NET)
                T1, C1, _, _ = s.state()
                Tw, Cw = so.plain_advance(T0, C0, S, occ, s.params())
                out.append((so.rel_l2(T1.astype(np.float64) - T0, Tw - T0),
                            so.rel_l2(C1.astype(np.float64) - C0, Cw - C0)))
            return np.mean(out, axis=0)
        e0 = err()
        loss = s.train(256, epochs=3, lr=1e-3, seed=5)
        e1 = err()
        assert np.isfinite(loss) and loss < 0.2
        assert e1[0] < 0.5 and e1[1] < 0.5 and e1[0] < 0.5 * e0[0], (e0, e1, loss)
