"""The schedule autotuner's choice and cache (host logic; the solver is a
stand-in that records the schedules applied to it)."""
import json

import numpy as np


class FakeSolver:
    def __init__(self, cost):
        self.cost, self.applied, self.current, self.solves = cost, [], None, 0

    def tune(self, gk=-1, trail_wg=0, persist_solve=-1, solve_wg=0, trail_tile=0, lookahead=-1):
        self.current = dict(gk=gk, trail_wg=trail_wg, solve_wg=solve_wg, trail_tile=trail_tile, lookahead=lookahead)
        self.applied.append(self.current)

    def solve(self, slots, vals, b, x):
        import time
        self.solves += 1
        time.sleep(self.cost(self.current))


def test_apply_fills_knobs_missing_from_old_entries():
    from fdmnes_amd import autotune
    s = FakeSolver(lambda c: 0)
    autotune.apply(s, {"gk": 12, "trail_wg": 768, "solve_wg": 512})  # an entry written before the newer knobs
    assert s.current == dict(gk=12, trail_wg=768, solve_wg=512, trail_tile=64, lookahead=0)


def test_tune_picks_fastest_and_caches_every_knob(tmp_path, monkeypatch):
    import torch

    from fdmnes_amd import autotune
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "cache.json"))
    monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)
    cands = [{"gk": 8, "trail_wg": 1024, "solve_wg": 512, "trail_tile": 64},
             {"gk": 20, "trail_wg": 1 << 24, "solve_wg": 512, "trail_tile": 64, "lookahead": 1},
             {"gk": 12, "trail_wg": 512, "solve_wg": 512, "trail_tile": 128}]
    lb1, lb2 = np.arange(1, 9, dtype=np.int32), np.arange(1, 9, dtype=np.int32)
    # the second candidate is fastest; the two-tile one is skipped for FP64
    s = FakeSolver(lambda c: 0.001 if c["gk"] == 20 else 0.02)
    ent = autotune.tune(s, [0], None, None, None, lb1, lb2, 0, candidates=cands)
    assert ent["best"] == {"gk": 20, "trail_wg": 1 << 24, "solve_wg": 512, "trail_tile": 64, "lookahead": 1}
    assert [a["trail_tile"] for a in s.applied[:-1]] == [64, 64]
    assert s.current["gk"] == 20 and s.solves == 4
    saved = json.load(open(tmp_path / "cache.json"))
    assert list(saved.values())[0]["best"] == ent["best"]
    # a second structure-identical call takes the cache and times nothing
    s2 = FakeSolver(lambda c: 1.0)
    ent2 = autotune.tune(s2, [0], None, None, None, lb1, lb2, 0, candidates=cands)
    assert ent2["best"] == ent["best"] and s2.solves == 0 and s2.current["gk"] == 20
