"""The persisted table of tuned solve-product shapes
(jf_2026_05_19_amd/tuning.py) on the CPU with a stand-in model: shape keys
bucket the front count by powers of two and the sizes by 64; a table
round-trips through its JSON file; apply sets stored shapes for known
classes and defaults elsewhere."""
import numpy as np

from jf_2026_05_19_amd import tuning


class FakeModel:
    def __init__(self, heights):
        self.h = np.asarray(heights, np.int64)
        self.cfg = np.zeros((len(heights), 2, 4), np.int32)

    def solve_heights(self):
        return self.h

    def solve_cfg(self):
        return self.cfg

    def set_solve_cfg(self, cfg):
        self.cfg = np.asarray(cfg, np.int32).reshape(-1, 2, 4)


def test_keys_bucket_counts_and_sizes():
    keys = tuning.shape_keys([[566, 336, 180, 0], [5, 64, 65, 0], [1, 1, 1, 0]])
    assert keys[0] == ("fwd:1024:6:3", "bwd:1024:6:3")
    assert keys[1] == ("fwd:8:1:2", "bwd:8:1:2")
    assert keys[2] == ("fwd:1:1:1", "bwd:1:1:1")


def test_table_roundtrip_and_apply(tmp_path):
    dm = FakeModel([[566, 336, 180, 0], [256, 264, 36, 0]])
    dm.cfg[:] = [[[384, 128, 8, 2], [256, 256, 16, 1]], [[128, 128, 32, 1], [384, 384, 8, 4]]]
    # tune() merges whatever the device autotuner left on the model
    dm.autotune = lambda reps: (1.0, 0.9, 10, 2)
    table, rep = tuning.tune(dm, {})
    assert rep == (1.0, 0.9, 10, 2) and len(table) == 4
    path = str(tmp_path / "t.json")
    tuning.save(path, table)
    assert tuning.load(path) == table
    fresh = FakeModel([[566, 336, 180, 0], [256, 264, 36, 0], [3, 900, 300, 0]])
    assert tuning.apply(fresh, tuning.load(path)) == 4
    assert np.array_equal(fresh.cfg[:2], dm.cfg)
    assert np.array_equal(fresh.cfg[2], [list(tuning.DEFAULT)] * 2)
    assert tuning.load(str(tmp_path / "missing.json")) == {}


def test_three_value_table_loads_with_one_k_group(tmp_path):
    path = tmp_path / "old.json"
    path.write_text('{"version": 1, "shapes": {"fwd:8:1:2": [256, 128, 16]}}')
    assert tuning.load(str(path)) == {"fwd:8:1:2": (256, 128, 16, 1)}
