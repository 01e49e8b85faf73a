"""Host logic of the launch-shape autotuner (coordinate descent, budgets,
inadmissible shapes), on a stand-in handle with a synthetic cost model."""
import pytest

from brian_amd import autotune


class FakeHandle:
    def __init__(self, wmax=11, bad=None):
        self.cfg = {"window": wmax, "neurons_per_thread": 4, "wave_max": 4, "row_split": 64,
                    "window_max": wmax}
        self.steps = 0
        self.bad = bad or (lambda cfg: False)

    def get_tuning(self, k):
        return self.cfg[k]

    def set_tuning(self, k, v):
        old = self.cfg[k]
        self.cfg[k] = v
        if self.bad(self.cfg):
            self.cfg[k] = old
            raise RuntimeError("inadmissible")

    def run(self, n):
        self.steps += n

    def sync(self):
        pass

    def cost(self):
        c = self.cfg
        return (abs(c["neurons_per_thread"] - 2) + abs(c["window"] - 8) * 0.5
                + (0 if c["wave_max"] == 16 else 1) + (0 if c["row_split"] == 16 else 0.7) + 10)


@pytest.fixture
def fake_timing(monkeypatch):
    def seg(h, steps, reps):
        h.steps += steps * reps
        return h.cost() * steps
    monkeypatch.setattr(autotune, "_time_segment", seg)


def test_finds_separable_optimum(fake_timing):
    h = FakeHandle()
    log = []
    best, used = autotune.tune(h, trial_steps=100, log=log)
    assert best == {"neurons_per_thread": 2, "window": 8, "wave_max": 16, "row_split": 16}
    assert {k: h.cfg[k] for k in best} == best      # handle left on the best shape
    assert used == h.steps
    assert len(log) > 10


def test_step_budget_respected(fake_timing):
    h = FakeHandle()
    best, used = autotune.tune(h, trial_steps=100, max_steps=2000)
    assert used <= 2000 and used == h.steps
    best0, used0 = autotune.tune(FakeHandle(), trial_steps=100, max_steps=100)
    assert used0 == 0


def test_inadmissible_shapes_skipped(fake_timing):
    # window x neurons_per_thread above a bound is refused by the engine
    h = FakeHandle(bad=lambda c: c["window"] * c["neurons_per_thread"] > 16)
    best, _ = autotune.tune(h, trial_steps=100)
    assert best["window"] * best["neurons_per_thread"] <= 16


def test_window_candidates_within_ring_bound():
    c = autotune.candidates(FakeHandle(wmax=3))
    assert max(c["window"]) == 3 and min(c["window"]) == 1
