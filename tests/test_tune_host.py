"""Host-side logic of the workgroup-shape autotuner (no GPU needed): the
tuning table lookup and the forced-shape override."""
import json

import pytest


def test_shape_lookup(tmp_path, monkeypatch):
    from riptide_amd import tune
    table = {"batch": {"240-260": 0, "480-520": 2}, "periodogram": {"240-260": 1}}
    path = tmp_path / "tuning.json"
    path.write_text(json.dumps(table))
    monkeypatch.setenv("RIPTIDE_AMD_TUNING", str(path))
    tune._cache = None
    assert tune.shape_for("batch", 240) == 0
    assert tune.shape_for("batch", 250) == 0
    assert tune.shape_for("batch", 500) == 2        # nearest class lower bound
    assert tune.shape_for("periodogram", 480) == 1  # only class present
    assert tune.shape_for("unknown", 240) is None
    tune._cache = None


def test_no_table(tmp_path, monkeypatch):
    from riptide_amd import tune
    monkeypatch.setenv("RIPTIDE_AMD_TUNING", str(tmp_path / "missing.json"))
    tune._cache = None
    assert tune.shape_for("batch", 240) is None
    tune._cache = None


def test_forced_shape_wins(tmp_path, monkeypatch):
    from riptide_amd import _lib, tune
    try:
        _lib.lib()
    except Exception:
        pytest.skip("native library not built")
    path = tmp_path / "tuning.json"
    path.write_text(json.dumps({"batch": {"240-260": 0}}))
    monkeypatch.setenv("RIPTIDE_AMD_TUNING", str(path))
    tune._cache = None
    try:
        tune.set_shape(2)
        assert tune.shape_for("batch", 240) == 2
        with pytest.raises(ValueError):
            tune.set_shape(7)
    finally:
        tune.set_shape(None)
        tune._cache = None
    assert tune.shape_for("batch", 240) == 0
    tune._cache = None


def test_command_line_runs_package_module(monkeypatch):
    """`python -m riptide_amd.tune` must act on the package's module, whose
    forced shape libcpp reads (a separate `__main__` copy would force a shape
    nobody reads, and every one-series timing would use the table's shape)."""
    import runpy
    from riptide_amd import tune
    calls = []
    monkeypatch.setattr(tune, "main", lambda argv=None: calls.append(argv))
    runpy.run_module("riptide_amd.tune", run_name="__main__")
    assert calls == [None]
