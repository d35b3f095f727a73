"""The autotuner's table as JSON (tbfs_tune_save / tbfs_tune_load): a
hand-written file loads, saves back field for field, merges by key, and a
malformed file is refused without touching the table.  Host code only (no
GPU call)."""
import json

import pytest

import turbobfs_amd as tb

ENTRIES = [
    {"words": 0, "log2_n": 21, "log2_max_in_degree": 16, "log2_twice_mean_in_degree": 6,
     "grid_per_cu": 0, "batch_alpha": 0, "alpha": 10, "beta": 48, "stage_ratio": 0.5,
     "lb_grid": 512, "ms": 0.17, "default_ms": 0.18},
    {"words": 16, "log2_n": 19, "log2_max_in_degree": 14, "log2_twice_mean_in_degree": 6,
     "grid_per_cu": 8, "batch_alpha": 32, "alpha": 0, "beta": 0, "stage_ratio": 0,
     "lb_grid": 0, "ms": 9.5, "default_ms": 0},
]


@pytest.fixture
def clean_table():
    tb.tune_clear()
    yield
    tb.tune_clear()


def test_json_round_trip(tmp_path, clean_table):
    src = tmp_path / "in.json"
    src.write_text(json.dumps({"version": 1, "entries": ENTRIES}, indent=1))
    tb.tune_load(str(src))
    assert tb.tune_count() == 2
    out = tmp_path / "out.json"
    tb.tune_save(str(out))
    back = json.loads(out.read_text())["entries"]
    assert len(back) == 2
    for a, b in zip(ENTRIES, back):
        for k, v in a.items():
            assert b[k] == pytest.approx(v), k
    # loading again merges by key (no duplicates), a changed value replaces
    ENTRIES2 = [dict(ENTRIES[0], alpha=20)]
    src.write_text(json.dumps({"version": 1, "entries": ENTRIES2}))
    tb.tune_load(str(src))
    assert tb.tune_count() == 2
    tb.tune_save(str(out))
    back = json.loads(out.read_text())["entries"]
    assert [e["alpha"] for e in back if e["words"] == 0] == [20]


def test_malformed_file_refused(tmp_path, clean_table):
    bad = tmp_path / "bad.json"
    bad.write_text('{"version": 1, "entries": [{"words": 0, "log2_n": 3}]}')  # key fields missing
    with pytest.raises(tb.TbfsError):
        tb.tune_load(str(bad))
    assert tb.tune_count() == 0
    with pytest.raises(tb.TbfsError):
        tb.tune_load(str(tmp_path / "missing.json"))
