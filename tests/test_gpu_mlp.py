"""The warm-start network's HIP kernels (BF16 MFMA products, FP32 accumulation) against
the CPU oracle (oracle/mlp_oracle.py).  Layer-local checks feed the oracle the GPU's own
BF16 operands, so each product and epilogue is compared at FP32 rounding; the end-to-end
gradient is compared at the BF16 level; training must reduce the loss."""
import os
import sys

import numpy as np
import pytest

import multiphaselb_amd as M

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as MO  # noqa: E402

pytestmark = pytest.mark.gpu

WIDTHS = [9, 256, 256, 256, 128]
B = 64


def setup(seed=0):
    rng = np.random.default_rng(seed)
    net = M.Mlp(WIDTHS, B)
    parts = []
    for l in range(len(WIDTHS) - 1):
        lim = np.sqrt(6.0 / (WIDTHS[l] + WIDTHS[l + 1]))
        parts += [rng.uniform(-lim, lim, WIDTHS[l] * WIDTHS[l + 1]), rng.uniform(-0.1, 0.1, WIDTHS[l + 1])]
    p = np.concatenate(parts).astype(np.float32)
    net.set_params(p)
    x = rng.uniform(-1, 1, (B, WIDTHS[0])).astype(np.float32)
    t = rng.uniform(-1, 1, (B, WIDTHS[-1])).astype(np.float32)
    return net, p, x, t


def test_param_count_and_roundtrip():
    net, p, _, _ = setup()
    assert net.P == sum(WIDTHS[l] * WIDTHS[l + 1] + WIDTHS[l + 1] for l in range(len(WIDTHS) - 1))
    np.testing.assert_array_equal(net.get_params(), p)


def test_layer_products_match_oracle():
    """Each layer's forward product and backward products, from the GPU's own BF16
    operands, against float64 sums: agreement to FP32 accumulation rounding."""
    net, p, x, t = setup(1)
    loss, g = net.gradient(x, t)
    L = len(WIDTHS) - 1
    wp = [32] + WIDTHS[1:]
    lay = MO.layers(p.astype(np.float64), WIDTHS)
    a = [net.tensor("a", j, (B, wp[j])).astype(np.float64) for j in range(L)]
    np.testing.assert_array_equal(a[0][:, :9], MO.bf16(x))
    assert not a[0][:, 9:].any()
    goff = 0
    for l in range(L):
        W, b = lay[l]
        wb = net.tensor("wb", l, (wp[l], WIDTHS[l + 1])).astype(np.float64)
        np.testing.assert_array_equal(wb[:W.shape[0]], MO.bf16(W.astype(np.float32)))
        z = a[l] @ wb + b.astype(np.float32)
        if l < L - 1:
            gpu = net.tensor("a32", l + 1, (B, wp[l + 1]))
            err = np.max(np.abs(gpu - np.tanh(z)))
        else:
            gpu = net.tensor("y", 0, (B, WIDTHS[-1]))
            err = np.max(np.abs(gpu - z)) / np.max(np.abs(z))
        print("layer %d forward err %.2e" % (l, err), flush=True)
        assert err < 2e-5
        # weight gradient from the GPU's BF16 operands
        dzb = net.tensor("dzb", l + 1, (B, WIDTHS[l + 1])).astype(np.float64)
        dz32 = net.tensor("dz", l + 1, (B, WIDTHS[l + 1])).astype(np.float64)
        np.testing.assert_array_equal(dzb, MO.bf16(dz32.astype(np.float32)))
        nW = W.size
        gW = g[goff:goff + nW].reshape(W.shape)
        gb = g[goff + nW:goff + nW + W.shape[1]]
        goff += nW + W.shape[1]
        ref = (a[l].T @ dzb)[:W.shape[0]]
        e_w = np.max(np.abs(gW - ref)) / np.max(np.abs(ref))
        e_b = np.max(np.abs(gb - dz32.sum(axis=0))) / np.max(np.abs(gb))
        print("layer %d dW err %.2e db err %.2e" % (l, e_w, e_b), flush=True)
        assert e_w < 1e-5 and e_b < 1e-5
        if l > 0:
            a32 = net.tensor("a32", l, (B, wp[l])).astype(np.float64)
            ref_dz = (dzb @ wb.T) * (1.0 - a32 * a32)
            got = net.tensor("dz", l, (B, wp[l]))
            e_x = np.max(np.abs(got - ref_dz)) / np.max(np.abs(ref_dz))
            print("layer %d dz err %.2e" % (l, e_x), flush=True)
            assert e_x < 1e-5


def test_gradient_matches_oracle_end_to_end():
    net, p, x, t = setup(2)
    loss, g = net.gradient(x, t)
    l16, g16 = MO.loss_grad(p, WIDTHS, x.astype(np.float64), t.astype(np.float64), bf16_ops=True)
    l64, g64 = MO.loss_grad(p, WIDTHS, x.astype(np.float64), t.astype(np.float64), bf16_ops=False)
    e16 = np.linalg.norm(g - g16) / np.linalg.norm(g16)
    e64 = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    print("loss %.6f oracle %.6f; grad err vs BF16 oracle %.2e, vs float64 %.2e" % (loss, l16, e16, e64), flush=True)
    assert abs(loss - l16) < 1e-4 * l16
    assert e16 < 5e-3 and e64 < 3e-2


def test_adam_step_and_training_reduce_loss():
    """One Adam step equals the closed form from the GPU gradient; training on a smooth
    target drives the loss down by orders of magnitude."""
    net, p, x, t = setup(3)
    _, g = net.gradient(x, t)
    idx = np.arange(B, dtype=np.int32)
    hist = net.train(x, t, idx, lr=1e-3)
    p1 = net.get_params()
    ref, _, _ = MO.adam_step(p.astype(np.float64), g.astype(np.float64), 0, 0, 1, 1e-3)
    assert np.max(np.abs(p1 - ref)) < 2e-6
    rng = np.random.default_rng(4)
    X = rng.uniform(-1, 1, (1024, WIDTHS[0])).astype(np.float32)
    s = np.linspace(0, 1, WIDTHS[-1])
    T = (np.tanh(3 * X[:, :1] * (s - 0.5)) * X[:, 1:2] + 0.3 * X[:, 2:3]).astype(np.float32)
    net.set_params(p)
    l0, _ = net.gradient(X[:B], T[:B])
    idx = rng.integers(0, 1024, 600 * B).astype(np.int32)
    hist = net.train(X, T, idx, lr=2e-3)
    print("loss start %.3e, last 20 mean %.3e" % (l0, hist[-20:].mean()), flush=True)
    assert np.isfinite(hist).all() and hist[-20:].mean() < 0.02 * l0
