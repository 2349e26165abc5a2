# SPDX-License-Identifier: MIT
"""The learned warm start's MFMA kernels against the oracle's network
(same bf16 rounding points, FP32 sums in index order; the matrix unit sums
each 16-term block in its own order and a rounding of an activation may
then flip by one bf16 step, hence the tolerances stated below)."""
import sys

import numpy as np
import pytest

import _oracle as O
import pmamd

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(0, scale, pmamd.MLP_PARAMS).astype(np.float32)


def test_mfma_forward_matches_the_oracle():
    say("warm-start forward, 200 rows")
    rng = np.random.default_rng(1)
    p = random_params(2)
    x = rng.normal(0, 1, (200, 16)).astype(np.float32)
    y = pmamd.warmstart_forward(p, x)
    ref = O.mlp_forward(p, x, round_bf16=True)
    scale = np.abs(ref).max()
    # |dy| <= 2% of the output scale (bf16 activation flips), and the rows
    # agree as vectors
    assert np.abs(y - ref).max() <= 0.02 * scale, np.abs(y - ref).max() / scale
    assert np.corrcoef(y.ravel(), ref.ravel())[0, 1] > 0.9999


def test_mfma_gradients_match_the_oracle():
    say("warm-start gradients, 64 samples")
    rng = np.random.default_rng(3)
    p = random_params(4)
    samples = rng.normal(0, 1, (64, 18)).astype(np.float32)
    loss, g = pmamd.warmstart_gradients(p, samples)
    ref_loss, ref = O.mlp_gradients(p, samples, round_bf16=True)
    assert abs(loss - ref_loss) <= 0.01 * ref_loss
    for name, (a, b) in {"W1": (0, 1024), "b1": (1024, 1088), "W2": (1088, 5184), "b2": (5184, 5248),
                         "W3": (5248, 5376), "b3": (5376, 5378)}.items():
        ga, gr = g[a:b], ref[a:b]
        cos = float(np.dot(ga, gr) / (np.linalg.norm(ga) * np.linalg.norm(gr) + 1e-30))
        say(f"  {name}: cosine {cos:.6f}, max |diff| / max |g| {np.abs(ga - gr).max() / np.abs(gr).max():.2e}")
        # the operands are the same bf16 values on both sides; only FP32
        # summation order differs
        assert cos > 0.999999, name
        assert np.abs(ga - gr).max() <= 1e-4 * np.abs(gr).max(), name


def test_training_fits_a_teacher_network():
    """Targets from a fixed random network of the same shape: Adam on the
    GPU takes the loss below a tenth of the untrained network's."""
    say("warm-start training on a teacher network, 16384 samples")
    rng = np.random.default_rng(5)
    teacher = random_params(6, scale=0.25)
    x = rng.normal(0, 1, (16384, 16)).astype(np.float32)
    t = O.mlp_forward(teacher, x, round_bf16=False)
    samples = np.concatenate([x, t], axis=1).astype(np.float32)
    p0 = pmamd.warmstart_init(7)
    loss0, _ = O.mlp_gradients(p0, samples[:4096], round_bf16=False)
    p, losses = pmamd.warmstart_train(samples, epochs=8, learning_rate=2e-3, seed=7)
    loss1, _ = O.mlp_gradients(p, samples[:4096], round_bf16=False)
    say(f"  loss {loss0:.4g} -> {loss1:.4g}; per epoch {np.array2string(losses, precision=4)}")
    assert loss1 < 0.1 * loss0
    assert losses[-1] < losses[0]
