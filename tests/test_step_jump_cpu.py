"""The raycast reaches the first sample of a tile's depth range by jumping
through the FP32 sequence t, t + step, ... (t += step, as the reference's
march takes it, ray_caster.cu:56-92) instead of adding step by step
(advance_steps in meshhashing_amd/csrc/raycast.hip).  This restates that
procedure in numpy float32 / Python float64 and checks it against plain
step-by-step float32 addition, including steps that are exact rounding ties
on some binade's grid (where round-half-even alternates before it settles).
The GPU raycast parity tests check the kernel itself against the oracle.
"""
import math

import numpy as np

f32 = np.float32


def brute(t, step, target):
    t, step = f32(t), f32(step)
    while t < target:
        t = f32(t + step)
    return t


def advance(t, step, target):
    t, step, target = f32(t), f32(step), f32(target)
    while t < target:
        t1 = f32(t + step)
        if not t1 < target:
            return t1
        t2 = f32(t1 + step)
        if not t2 < target:
            return t2
        t3 = f32(t2 + step)
        if not t3 < target:
            return t3
        lo = 2.0 ** (math.frexp(float(t3))[1] - 1)  # t3 in [lo, 2 lo)
        if float(t1) < lo:
            t = t3
            continue
        top, u = 2.0 * lo, lo * 2.0 ** -23
        d = float(t3) - float(t2)
        room = (top - 2.0 * u - float(step) - float(t3)) / d
        if not room >= 1.0:
            t = t3
            continue
        k = min(math.floor(room) + 1.0, math.ceil((float(target) - float(t3)) / d))
        t = f32(float(t3) + k * d)
    return t


def test_jump_equals_stepwise_addition():
    rng = np.random.default_rng(3)
    steps = [0.01, 0.004, 0.0025, 0.005, 0.02]
    for i in range(3000):
        t0 = f32(rng.uniform(0.05, 0.5))
        step = f32(rng.choice(steps) * (rng.uniform(0.5, 2.0) if i % 3 else 1.0))
        target = f32(t0 + rng.uniform(0.0, 6.0))
        assert advance(t0, step, target) == brute(t0, step, target), (t0, step, target)


def test_jump_on_tie_steps():
    # step = (n + 1/2) ulps of some binade: every sum in that binade is a tie
    rng = np.random.default_rng(4)
    ties = 0
    for _ in range(1500):
        u = 2.0 ** (int(rng.integers(-3, 2)) - 23)
        step = f32((int(rng.integers(1000, 200000)) + 0.5) * u)
        t0 = f32(rng.uniform(0.05, 0.3))
        target = f32(t0 + rng.uniform(0.0, 4.0))
        assert advance(t0, step, target) == brute(t0, step, target), (t0, step, target)
        ties += 1
    assert ties == 1500
