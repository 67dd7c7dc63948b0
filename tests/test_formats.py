"""On-disk formats the reference's harness relies on (CPU; hipcc compiles,
nothing runs on a GPU): the .cache_bitgen run layout and the -r reuse of the
latest run (bitgen/from_icgrep.py:314-348), and the timing CSV schema
(bitgen/tool/timer.py:58-85, 164-172)."""
import os

import pytest

from bitgen_amd import config as cfg
from bitgen_amd import from_icgrep
from bitgen_amd.tool.timer import Timer


@pytest.fixture
def cache_root(tmp_path, monkeypatch):
    saved = dict(cfg.get_config_dict())
    cfg.reset_config()
    cfg.apply_hip_defaults()
    root = tmp_path / ".cache_bitgen"
    cfg.set_config("cache_dir", str(root))
    cfg.set_config("hsaco_cache_dir", str(tmp_path / "hsaco"))
    monkeypatch.chdir(tmp_path)
    yield root
    cfg.set_from_dict(saved)


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc"), reason="hipcc not installed")
def test_cache_layout_and_reuse(cache_root):
    groups = [["ab[0-9]{1,3}c"], ["hello", "x.*y"]]
    path = from_icgrep.compile_regex_group_list(groups)
    run_dir = os.path.dirname(os.path.dirname(os.path.dirname(path)))
    assert os.path.basename(run_dir).startswith("bitgen_code_")
    assert path.endswith(os.path.join("code", "kernels", "kernel_0.py"))
    kdir = os.path.dirname(path)
    assert open(os.path.join(kdir, "__init__.py")).read() == "import kernels.kernel_0 as kernel_0\n"
    assert sorted(f for f in os.listdir(kdir) if f.endswith(".hsaco")) == ["kernel_0.hsaco", "kernel_1.hsaco"]
    # a second compile in the same run goes to code_1
    path2 = from_icgrep.compile_regex_group_list(groups)
    assert os.path.join("code_1", "kernels") in path2
    # -r: a new process (fresh config) reuses the latest run's code without
    # regenerating it, although the cache root also holds other directories
    mtime = os.path.getmtime(path)
    for key in ("__cached_path", "__cached_module_path"):
        cfg.get_config_dict().pop(key, None)
    cfg.set_config("use_cached_code", True)
    os.makedirs(cache_root / "tune", exist_ok=True)
    os.makedirs(cache_root / "zzz_other", exist_ok=True)
    again = from_icgrep.compile_regex_group_list(groups)
    assert again == path and os.path.getmtime(path) == mtime


def test_reuse_without_a_cached_run_fails_clearly(cache_root):
    os.makedirs(cache_root / "hsaco", exist_ok=True)
    cfg.set_config("use_cached_code", True)
    with pytest.raises(FileNotFoundError, match="bitgen_code_"):
        from_icgrep.init_cached_path()


def test_timing_csv_schema(tmp_path):
    import pandas as pd

    t = Timer()
    t.sync = lambda: None  # no device to synchronise
    t.set_group("regex_group_0")
    t.set_exe("bitgen")
    t.set_app("snort")
    t.add_timing("transpose", 0.5)
    t.add_timing("run_regex", 2.0)
    t.add_timing("run_regex", 4.0)
    for k, v in (("input_size", 0.0095), ("throughput", 3.2), ("count", 16), ("ref_count", 16), ("check", True)):
        t.append_data("run_regex", k, v)
    path = t.save_to_file(str(tmp_path / "raw_results" / "ac"))
    df = pd.read_csv(path)
    assert list(df.columns[:7]) == ["group", "exe", "app", "name", "exec_num", "duration", "avg_duration"]
    row = df[df["name"] == "run_regex"].iloc[0]
    assert row["exec_num"] == 2 and row["duration"] == 6.0 and row["avg_duration"] == 3.0
    assert row["count"] == 16 and row["ref_count"] == 16 and bool(row["check"])
    assert set(["input_size", "throughput", "count", "ref_count", "check"]) <= set(df.columns)
