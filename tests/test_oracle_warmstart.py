# SPDX-License-Identifier: MIT
"""The learned warm start's network on the oracle (no reference
counterpart): its forward pass and mean-squared-error gradients, without the
bf16 rounding points, equal an independent float64 restatement in numpy,
whose gradients are themselves pinned by central finite differences."""
import numpy as np

import _oracle as O

IN, HID, OUT = 16, 64, 2
SIZES = [HID * IN, HID, HID * HID, HID, OUT * HID, OUT]


def unpack(p):
    out, k = [], 0
    for n in SIZES:
        out.append(np.asarray(p[k:k + n], np.float64))
        k += n
    w1, b1, w2, b2, w3, b3 = out
    return w1.reshape(HID, IN), b1, w2.reshape(HID, HID), b2, w3.reshape(OUT, HID), b3


def np_loss_grads(p, samples):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x, t = samples[:, :IN].astype(np.float64), samples[:, IN:].astype(np.float64)
    z1 = x @ w1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ w3.T + b3
    n = len(x)
    loss = np.sum((y - t) ** 2) / (n * OUT)
    dy = 2.0 * (y - t) / (n * OUT)
    gw3, gb3 = dy.T @ h2, dy.sum(0)
    dz2 = (dy @ w3) * (z2 > 0)
    gw2, gb2 = dz2.T @ h1, dz2.sum(0)
    dz1 = (dz2 @ w2) * (z1 > 0)
    gw1, gb1 = dz1.T @ x, dz1.sum(0)
    return loss, np.concatenate([gw1.ravel(), gb1, gw2.ravel(), gb2, gw3.ravel(), gb3]), y


def random_case(seed, n=48):
    rng = np.random.default_rng(seed)
    p = np.concatenate([rng.normal(0, 0.3, s) for s in SIZES]).astype(np.float32)
    samples = rng.normal(0, 1, (n, IN + OUT)).astype(np.float32)
    return p, samples


def test_numpy_gradients_match_finite_differences():
    p, samples = random_case(1, n=8)
    _, g, _ = np_loss_grads(p.astype(np.float64), samples)
    rng = np.random.default_rng(2)
    for k in rng.choice(len(p), 40, replace=False):
        h = 1e-6
        pp, pm = p.astype(np.float64).copy(), p.astype(np.float64).copy()
        pp[k] += h
        pm[k] -= h
        fd = (np_loss_grads(pp, samples)[0] - np_loss_grads(pm, samples)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_oracle_forward_and_gradients_equal_the_float64_restatement():
    for seed in range(3):
        p, samples = random_case(seed)
        loss, g = O.mlp_gradients(p, samples, round_bf16=False)
        ref_loss, ref_g, ref_y = np_loss_grads(p, samples)
        y = O.mlp_forward(p, samples[:, :IN], round_bf16=False)
        np.testing.assert_allclose(y, ref_y, rtol=1e-5, atol=1e-5)
        assert abs(loss - ref_loss) <= 1e-5 * ref_loss
        np.testing.assert_allclose(g, ref_g, rtol=1e-4, atol=1e-6 * np.abs(ref_g).max())


def test_bf16_rounding_points_move_results_by_bf16_amounts():
    p, samples = random_case(5)
    _, g32 = O.mlp_gradients(p, samples, round_bf16=False)
    _, g16 = O.mlp_gradients(p, samples, round_bf16=True)
    rel = np.abs(g16 - g32).max() / np.abs(g32).max()
    assert 0.0 < rel < 0.05, rel
