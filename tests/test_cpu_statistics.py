"""brian/tests/testinterface/test_statistics.py and the statistics that
judge fitting (brian/tools/statistics.py:15-30, :236-306), pinned by hand
cases, a brute-force double loop and the expectation for independent
Poisson trains."""
import numpy as np
import pytest

import brian_amd as b


def test_empty_trains_give_nan():
    """test_statistics.py:4-11 (the rate and CV lines)."""
    assert b.firing_rate([]) is b.NaN
    assert b.CV([]) is b.NaN


def test_cv_small_and_regular_trains():
    """One spike has no interval (NaN); two spikes have one interval and no
    spread (0); a regular train has CV 0 and rate 1/period; ISIs 1, 3 give
    std 1 over mean 2."""
    assert np.isnan(b.CV([0.5]))
    assert b.CV([0.1, 0.3]) == 0.0
    t = np.arange(100) * 0.01
    assert abs(b.firing_rate(t) - 100.0) < 1e-9
    assert b.CV(t) < 1e-9
    assert abs(b.CV([0.0, 1.0, 4.0]) - 0.5) < 1e-15


def test_gamma_hand_case():
    """Targets are matched by their nearest source: only 1.001 finds one
    (1.002) within 4 ms, although two sources (1.0 and 1.002) lie within 4 ms
    of that target."""
    src = [1.0, 1.002, 3.0]
    tgt = [1.001, 2.0, 4.0]
    assert b.gamma_factor(src, tgt, 0.004, normalize=False) == 1
    assert b.gamma_factor(src, tgt, 0.004, normalize=False, dt=1e-4) == 1
    # the other direction matches two sources to the one close target
    assert b.gamma_factor(tgt, src, 0.004, normalize=False) == 2
    # a single source spike: every target within delta of it
    assert b.gamma_factor([2.0], [1.997, 2.003, 2.01], 0.004, normalize=False) == 2
    assert b.gamma_factor([], tgt, 0.004) == 0
    assert b.gamma_factor(src, [], 0.004) == 0


def _brute(source, target, delta):
    n = 0
    for t in target:
        n += min(abs(t - s) for s in source) <= delta
    return n


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gamma_count_brute_force(seed):
    """The count equals a double loop: targets whose nearest source is within
    delta, in seconds and on the step grid."""
    rng = np.random.default_rng(seed)
    src = np.sort(rng.random(120) * 2.0)
    tgt = np.sort(rng.random(90) * 2.0)
    assert b.gamma_factor(src, tgt, 0.004, normalize=False) == _brute(src, tgt, 0.004)
    dt = 1e-4
    si, ti = np.rint(src / dt).astype(int), np.rint(tgt / dt).astype(int)
    assert b.gamma_factor(src, tgt, 0.004, normalize=False, dt=dt) == _brute(si, ti, 40)
    assert b.gamma_factor(src[:1], tgt, 0.004, normalize=False) == _brute(src[:1], tgt, 0.004)


def test_gamma_identical_trains_is_one():
    t = np.cumsum(np.random.default_rng(7).exponential(0.05, 400))
    assert abs(b.gamma_factor(t, t, 0.004) - 1.0) < 1e-12
    assert abs(b.gamma_factor(t, t, 0.004, dt=1e-4) - 1.0) < 1e-12


@pytest.mark.parametrize("dt", [None, 1e-4])
def test_gamma_independent_poisson_near_zero(dt):
    """For independent Poisson trains at rate r the expected count is
    Nt (1 - exp(-2 delta r)), just below the normaliser's 2 delta r Nt: at
    10 Hz and 4 ms E[gamma] = -0.0034.  A rate taken per step or a window of
    delta instead of 2 delta moves gamma by +0.087 or +0.043."""
    rng = np.random.default_rng(11)
    g = []
    for _ in range(8):
        src = np.cumsum(rng.exponential(0.1, 3000))
        tgt = np.cumsum(rng.exponential(0.1, 3000))
        g.append(b.gamma_factor(src, tgt, 0.004, dt=dt))
    assert abs(np.mean(g) + 0.0034) < 0.01, g
