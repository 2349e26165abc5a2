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


def test_learned_penalty_warm_start_takes_fewer_iterations():
    """Logged from plain runs of two 4096-body piles, trained on the GPU,
    then on a third pile in to-residual mode (RMS 1e-4, cap 64): the
    predicted penalty takes fewer pose-phase iterations per step than the
    plain warm start (the gain this test shows), with few fallbacks."""
    import scenarios as S
    say("learned warm start: log, train, compare (4096-body piles)")
    logs = []
    for seed in (31, 32):
        w = S.pile(side=16, seed=seed).build(pmamd.World(profile=0, base_iterations=10))
        w.set_contact_log(2000000)
        w.step(200)
        log = w.contact_log()
        w.close()
        logs.append(log[np.random.default_rng(seed).choice(len(log), size=min(300000, len(log)), replace=False)])
    samples = np.concatenate(logs)
    params, loss = pmamd.warmstart_train(samples, epochs=6, learning_rate=1e-3, seed=3)
    say(f"  {len(samples)} samples, loss per epoch {np.array2string(loss, precision=4)}")
    scene = S.pile(side=16, seed=33)
    result = {}
    for mode in ("plain", "learned"):
        w = scene.build(pmamd.World(profile=0, base_iterations=64, residual_tolerance=1e-4))
        if mode == "learned":
            w.set_warmstart(params)
        w.step(30)
        its = []
        for _ in range(150):
            w.step(1)
            its.append(int(w.residuals()[1][0]))
        result[mode] = (float(np.mean(its)), w.warmstart_stats())
        w.close()
    say(f"  pose iterations per step: plain {result['plain'][0]:.2f}, learned {result['learned'][0]:.2f}; "
        f"learned (steps, fallbacks) {result['learned'][1]}")
    assert result["learned"][0] < result["plain"][0]
    steps, falls = result["learned"][1]
    assert falls <= 0.1 * steps


def test_fallback_reproduces_the_plain_step():
    """Two worlds step a 4096-body pile identically with the plain warm start
    into contact (40 steps); then one switches to the learned warm start with
    a residual tolerance that 3 iterations do not reach. Each step that falls
    back is re-run from the restored velocity, inputs, kinematic flags and
    contact-cache parity and must equal the plain world's step bit for bit
    (states and per-update residuals). A step the learned warm start passes
    must end within the tolerance; the worlds part there."""
    import scenarios as S
    say("learned warm start: fallback steps equal the plain step (4096-body pile)")
    tol = 1e-9
    params = pmamd.warmstart_init(7)
    params[pmamd.MLP_PARAMS - 1] = 0.3   # log10(k/k0) within [0, 0.3]
    scene = S.pile(side=16, seed=41)
    plain, learned = [scene.build(pmamd.World(profile=0, base_iterations=3, residual_tolerance=tol))
                      for _ in range(2)]
    for w in (plain, learned):
        w.step(40)
    assert plain.states().tobytes() == learned.states().tobytes()
    learned.set_warmstart(params)
    compared = 0
    for step in range(40):
        plain.step(1)
        before = learned.warmstart_stats()[1]
        learned.step(1)
        rb = learned.residuals()
        if learned.warmstart_stats()[1] == before:
            last = [rb[0][ph][rb[1][ph] - 1] for ph in range(2) if rb[1][ph]]
            assert all(x <= tol for x in last), f"step {step} passed with residuals {last}"
            break
        ra = plain.residuals()
        assert plain.states().tobytes() == learned.states().tobytes(), f"states differ at step {step}"
        assert ra[0].tobytes() == rb[0].tobytes() and np.array_equal(ra[1], rb[1])
        compared += 1
    say(f"  fallback steps equal to the plain step: {compared}; (steps, fallbacks) {learned.warmstart_stats()}")
    assert compared >= 20
    for w in (plain, learned):
        w.close()
