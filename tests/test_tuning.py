"""The autotuning table behind the automatic tuning (rnpx_tune.cpp): the
JSON reader, nearest-shape lookup, validation and the default-location
load.  Host logic only: no GPU needed."""
import json
import os

import pytest

import rnp2_amd as R

TABLE = {"version": 1, "device": "test", "entries": [
    {"kind": "factor", "dtype": "d", "m": 8192, "n": 8192, "nb": 128, "ksplit": 8, "ms": 28.0},
    {"kind": "factor", "dtype": "d", "m": 32768, "n": 4096, "nb": 64, "ksplit": 0, "ms": 37.0},
    {"kind": "factor", "dtype": "z", "m": 4096, "n": 4096, "nb": 32, "ksplit": 4, "panel_cus": 32, "ms": 70.0},
    {"kind": "batched", "dtype": "d", "lds": 81920, "blocked_min": 17, "ms": 8.0}],
    "candidates": {"factor d 8192x8192": [{"nb": 32, "ksplit": 0, "ms": 90.0}]}}


def write(tmp_path, doc, name="t.json"):
    p = tmp_path / name
    p.write_text(doc if isinstance(doc, str) else json.dumps(doc))
    return str(p)


@pytest.fixture(autouse=True)
def restore_default():
    yield
    R.lib().rnpx_tuning_load(None)


def test_load_and_nearest_lookup(tmp_path):
    assert R.tuning_load(write(tmp_path, TABLE)) == 4
    assert R.tuning_lookup_factor("d", 8192, 8192) == (128, 8)
    assert R.tuning_lookup_factor("d", 6000, 7000) == (128, 8)          # within 2x of 8192^2
    assert R.tuning_lookup_factor("d", 30000, 4096) == (64, 0)          # nearer the tall entry
    assert R.tuning_lookup_factor("d", 2000, 2000) is None              # more than 2x from every entry
    assert R.tuning_lookup_factor("z", 4096, 3000) == (32, 4)
    assert R.tuning_lookup_factor("z", 8192, 8192) == (32, 4)           # 2x in each: still in range
    assert R.tuning_lookup_batched("d") == (81920, 17)
    assert R.tuning_lookup_batched("z") is None
    # the optional panel stream CU count: given for the complex entry only
    assert R.tuning_lookup_panel_cus("z", 4096, 4096) == 32
    assert R.tuning_lookup_panel_cus("d", 8192, 8192) == 0
    assert R.tuning_lookup_panel_cus("d", 2000, 2000) is None


@pytest.mark.parametrize("bad", [
    "{", '{"entries": [{"kind": "factor", "dtype": "d", "m": 8, "n": 8, "nb": 48, "ksplit": 0}]}',   # width not built
    '{"entries": [{"kind": "factor", "dtype": "q", "m": 8, "n": 8, "nb": 32, "ksplit": 0}]}',        # element type
    '{"entries": [{"kind": "batched", "dtype": "d", "lds": 10, "blocked_min": 3}]}',                 # LDS out of range
    '{"entries": [{"kind": "factor", "dtype": "d", "m": 8, "n": 8, "nb": 32, "ksplit": 0, "panel_cus": -2}]}',  # CUs
    '{"entries": [{"kind": "factor" "dtype": "d"}]}',                                                # syntax
    '{"entries": []} trailing'])
def test_malformed_tables_are_refused_and_keep_the_old_one(tmp_path, bad):
    R.tuning_load(write(tmp_path, TABLE))
    st = R.lib().rnpx_tuning_load(write(tmp_path, bad, "bad.json").encode())
    assert st == R.RNPX_ERR_ARG
    assert R.tuning_lookup_factor("d", 8192, 8192) == (128, 8)


def test_missing_file(tmp_path):
    assert R.lib().rnpx_tuning_load(str(tmp_path / "absent.json").encode()) == R.RNPX_ERR_UNSUPPORTED


def test_committed_table_is_valid():
    """The table shipped beside the library loads and covers the benchmarked
    factorization shapes and the batched configuration."""
    path = os.path.join(os.path.dirname(R.__file__), "rnpx_tuning.json")
    if not os.path.exists(path):
        pytest.skip("no autotuned table committed yet")
    assert R.tuning_load(path) > 0
    assert R.tuning_lookup_factor("d", 8192, 8192) is not None
    assert R.tuning_lookup_batched("d") is not None


def test_pick_factor_prefers_shared_cus_within_margin():
    """The autotuner keeps the CUs shared unless a partition wins by more
    than the margin."""
    from rnp2_amd.autotune import pick_factor
    res = [{"nb": 128, "ksplit": 8, "panel_cus": 0, "ms": 28.40},
           {"nb": 128, "ksplit": 8, "panel_cus": 64, "ms": 28.25},
           {"nb": 64, "ksplit": 0, "panel_cus": 0, "ms": 29.90}]
    assert pick_factor(res)["panel_cus"] == 0                      # 0.5% slower: shared kept
    res[1]["ms"] = 27.50
    assert pick_factor(res) == res[1]                              # 3% faster: partition taken
    assert pick_factor(res, margin=0.05)["panel_cus"] == 0         # within a 5% margin again
    only = [{"nb": 32, "ksplit": 16, "panel_cus": 64, "ms": 26.2}, {"nb": 32, "ksplit": 4, "panel_cus": 32, "ms": 27.0}]
    assert pick_factor(only) == only[0]                            # no shared candidate: the fastest


def test_written_table_loads_back(tmp_path, monkeypatch):
    """A table written by the autotuner is read back by the library's loader
    with the same widths, splits, panel CUs and batched settings."""
    import rnp2_amd.autotune as T
    monkeypatch.setattr(T, "device_name", lambda: "test device")
    entries = [{"kind": "factor", "dtype": "d", "m": 8192, "n": 8192, "nb": 128, "ksplit": 8, "panel_cus": 0, "ms": 28.5},
               {"kind": "factor", "dtype": "z", "m": 4096, "n": 4096, "nb": 32, "ksplit": 16, "panel_cus": 64,
                "ms": 26.2},
               {"kind": "batched", "dtype": "d", "lds": 20480, "blocked_min": 33, "ms": 4.8}]
    path = T.write_table(entries, path=str(tmp_path / "t.json"), extra={"candidates": {"x": [{"ms": 1.0}]}})
    doc = json.loads(open(path).read())
    assert doc["device"] == "test device" and doc["entries"] == entries
    assert R.tuning_load(path) == 3
    assert R.tuning_lookup_factor("d", 8192, 8192) == (128, 8)
    assert R.tuning_lookup_factor("z", 4096, 4096) == (32, 16)
    assert R.tuning_lookup_panel_cus("z", 4096, 4096) == 64
    assert R.tuning_lookup_panel_cus("d", 8192, 8192) == 0
    assert R.tuning_lookup_batched("d") == (20480, 33)
