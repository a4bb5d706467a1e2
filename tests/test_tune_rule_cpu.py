"""The online marching-cubes tuner's shape rule (meshhashing_amd/csrc/mh_tune.h),
host code compiled with g++ and run on constructed trial times: the trial
order, the paired comparison with the neighbouring shape-0 trials, the
worst-ratio rule and its 0.97 threshold, the median default time."""
import math
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SHAPES = 9  # MH_NUM_MC_SHAPES
ROUND = 2 * (SHAPES - 1)


@pytest.fixture(scope="module")
def rule(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = tmp_path_factory.mktemp("tune") / "tune_rule"
    subprocess.run([gxx, "-std=c++17", "-O1", f"-I{ROOT / 'meshhashing_amd' / 'csrc'}",
                    str(ROOT / "tests" / "tune_rule_main.cpp"), "-o", str(exe)], check=True)

    def run(tp):
        text = f"{SHAPES} {len(tp)} " + " ".join("nan" if (isinstance(t, float) and math.isnan(t)) else repr(t)
                                                  for t in tp)
        out = subprocess.run([str(exe)], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
        shape, dflt, best = out[0].split()
        order = [int(x) for x in out[1].split()]
        return int(shape), float(dflt), float(best), order
    return run


def trials(reps=1, base=1.0):
    return [base] * (reps * ROUND + 1)


def cand_trials(c, reps=1):
    """Trial indices at which candidate c runs."""
    return [2 * (r * (SHAPES - 1) + c - 1) + 1 for r in range(reps)]


def test_trial_order(rule):
    _, _, _, order = rule(trials(2))
    assert order[0::2] == [0] * (ROUND + 1)
    assert order[1::2] == list(range(1, SHAPES)) * 2
    for c in range(1, SHAPES):
        assert [order[p] for p in cand_trials(c, 2)] == [c, c]


def test_equal_times_keep_shape_0(rule):
    assert rule(trials())[:3] == (0, 1.0, 1.0)


def test_faster_candidate_wins_and_threshold(rule):
    tp = trials()
    tp[cand_trials(3)[0]] = 0.9
    shape, dflt, best, _ = rule(tp)
    assert (shape, dflt) == (3, 1.0) and best == pytest.approx(0.9)
    tp[cand_trials(3)[0]] = 0.98  # not 3 % faster
    assert rule(tp)[0] == 0


def test_least_worst_ratio_of_several(rule):
    tp = trials()
    tp[cand_trials(2)[0]] = 0.95
    tp[cand_trials(6)[0]] = 0.90
    assert rule(tp)[0] == 6


def test_first_frame_outlier_does_not_pick_its_neighbour(rule):
    # the first (shape-0) trial costs five times a later one; candidate 1,
    # just after it, is no faster than the shape-0 trial on its other side
    tp = trials()
    tp[0] = 5.0
    shape, dflt, _, _ = rule(tp)
    assert shape == 0 and dflt == 1.0  # the median ignores the outlier


def test_worst_ratio_over_rounds(rule):
    tp = trials(2)
    a, b = cand_trials(4, 2)
    tp[a] = 0.9  # faster in one round only
    assert rule(tp)[0] == 0
    tp[b] = 0.92
    shape, _, best, _ = rule(tp)
    assert shape == 4 and best == pytest.approx(0.92)


def test_trial_without_time_rules_candidate_out(rule):
    tp = trials()
    tp[cand_trials(5)[0]] = 0.8
    tp[cand_trials(2)[0]] = float("nan")
    assert rule(tp)[0] == 5
    tp[cand_trials(5)[0]] = float("nan")
    assert rule(tp)[0] == 0


def test_drift_is_compared_locally(rule):
    # times drift down by 2 % a frame; a candidate no faster than its
    # neighbours stays out although it beats the early shape-0 trials
    tp = [1.0 * 0.98 ** p for p in range(ROUND + 1)]
    assert rule(tp)[0] == 0
    p = cand_trials(7)[0]
    tp[p] = 0.96 * min(tp[p - 1], tp[p + 1])
    assert rule(tp)[0] == 7
