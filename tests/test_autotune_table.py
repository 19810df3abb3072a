"""The autotuner's table: JSON round trip with register budgets, workgroups per
long chain and width multipliers, and the C entry point each entry goes to
(a recording stand-in for the library; no device)."""
from bdsim_amd import autotune


class Recorder:
    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def call(*args):
            self.calls.append((name,) + tuple(args))
            return 0
        return call


def test_table_round_trip(tmp_path):
    table = {64: 4, 128: 1, 256: 4, autotune.SPLIT_KEY: 8, autotune.width_key(128): 4,
             autotune.width_key(512): 2}
    p = tmp_path / "tune.json"
    autotune.save(table, p)
    assert autotune.load(p) == table


def test_apply_routes_each_entry():
    rec = Recorder()
    autotune.apply({64: 2, autotune.SPLIT_KEY: 4, autotune.width_key(256): 2}, rec)
    assert ("bdsim_tune_set", 64, 2) in rec.calls
    assert ("bdsim_tune_split", 4) in rec.calls
    assert ("bdsim_tune_width", 256, 2) in rec.calls
    rec = Recorder()
    autotune.reset(rec)
    names = {c[0] for c in rec.calls}
    assert names == {"bdsim_tune_set", "bdsim_tune_split", "bdsim_tune_width"}
    assert all(c[-1] == 0 for c in rec.calls)


def test_width_candidates_stay_within_1024_threads():
    for b, ms in autotune.WIDTH_CANDIDATES.items():
        assert all(b * m <= 1024 for m in ms) and ms[0] == 1
