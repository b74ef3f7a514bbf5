"""CPU checks of the autotuner's search and cache with a stand-in context
(no GPU): it keeps the fastest configuration, refuses a configuration that
changes the traversed-edge counts, and reuses its cache."""
import numpy as np
import pytest

from graph500_amd import autotune


class FakeCtx:
    """Batch time is a fixed function of the tuning; optionally one knob
    value corrupts the edge counts."""

    def __init__(self, bad=None, bad_pairs=None, grid_gain=0.5):
        self.t = {"alpha": 16.0, "pull_grid": 8192, "sweep_below": 0.25, "sweep_grid": 4096, "persistent": 0,
                  "persist_levels": 16, "persist_grid": 0, "persist_waves": 8, "pull_filter": 0, "beta": 50.0,
                  "compact_below": 0.0, "filter_from": 0.01, "pull_unroll": 0, "lane_rule": 0,
                  "mixed_push": 1}
        self.bad = bad
        self.bad_pairs = bad_pairs
        self.grid_gain = grid_gain
        self.runs = 0

    def graph_info(self):
        return {"scale_estimate": 10, "num_active": 900, "nnz": 30000, "reorder_mode": 2}

    def version(self):
        return "fake gfx950"

    def batch_tuning(self):
        return dict(self.t)

    def batch_set_tuning(self, **kw):
        self.t.update(kw)

    def run_bfs_batch(self, roots):
        self.runs += 1

    def batch_stats(self):
        t = self.t
        # the per-root rule wins only with its own alpha scale
        a = abs(t["alpha"] - 128.0) / 64 - 1 if t["lane_rule"] else abs(t["alpha"] - 2.0) / 8
        ms = 10.0 + a + (0 if t["beta"] == 20.0 else 1) + t["pull_filter"] \
            + (2 if t["persistent"] else 0) + (0 if t["pull_grid"] == 16384 else self.grid_gain)
        pairs = [3, 40, 500, 7]
        if self.bad_pairs and self.t.get(self.bad_pairs[0]) == self.bad_pairs[1]:
            pairs = [3, 41, 499, 7]  # a vertex given the wrong level: same reached set
        return {"device_ms": ms, "pairs": pairs}

    def batch_teps_edges(self):
        e = np.arange(4, dtype=np.int64)
        if self.bad and self.t.get(self.bad[0]) == self.bad[1]:
            e[0] += 1
        return e


def test_descent_finds_the_best_per_knob(tmp_path):
    c = FakeCtx()
    res = autotune.autotune(c, [1, 2], reps=1, cache=str(tmp_path / "t.json"))
    assert res["tuning"]["lane_rule"] == 1 and res["tuning"]["alpha"] == 128.0
    assert res["tuning"]["beta"] == 20.0
    assert res["tuning"]["pull_filter"] == 0
    assert res["tuning"]["persistent"] == 0
    assert res["tuning"]["pull_grid"] == 16384
    assert c.batch_tuning() == res["tuning"]


def test_cache_is_reused(tmp_path):
    path = str(tmp_path / "t.json")
    first = autotune.autotune(FakeCtx(), [1], reps=1, cache=path)
    c = FakeCtx()
    again = autotune.autotune(c, [1], cache=path)
    assert again["cached"] and again["tuning"] == first["tuning"]
    assert c.runs == 0 and c.batch_tuning() == first["tuning"]


def test_changed_results_are_refused(tmp_path):
    with pytest.raises(RuntimeError):
        autotune.autotune(FakeCtx(bad=("alpha", 8.0)), [1], reps=1, cache=str(tmp_path / "t.json"))


def test_wrong_levels_with_the_same_reached_set_are_refused(tmp_path):
    """A trial that moves a vertex to another level keeps every traversed-edge
    count; the per-level pair counts catch it."""
    with pytest.raises(RuntimeError, match="per level"):
        autotune.autotune(FakeCtx(bad_pairs=("pull_unroll", 3)), [1], reps=1, cache=str(tmp_path / "t.json"))


def test_union_rule_alpha_is_kept_when_it_wins(tmp_path):
    class UnionWins(FakeCtx):
        def batch_stats(self):
            st = super().batch_stats()
            return {"device_ms": st["device_ms"] + (5.0 if self.t["lane_rule"] else 0.0), "pairs": st["pairs"]}
    res = autotune.autotune(UnionWins(), [1], reps=1, cache=str(tmp_path / "t.json"))
    assert res["tuning"]["lane_rule"] == 0 and res["tuning"]["alpha"] == 2.0


def test_noise_sized_gains_do_not_move_the_tuning(tmp_path):
    """A knob value faster by less than min_gain (0.3 %) is not taken; one
    faster by more is."""
    small = FakeCtx(grid_gain=0.01)  # 16384 saves 0.01 ms of about 9 ms
    res = autotune.autotune(small, [1], reps=1, cache=str(tmp_path / "a.json"))
    assert res["tuning"]["pull_grid"] == 8192
    res = autotune.autotune(FakeCtx(grid_gain=0.01), [1], reps=1, cache=str(tmp_path / "b.json"), min_gain=0.0)
    assert res["tuning"]["pull_grid"] == 16384
