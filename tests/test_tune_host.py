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


def test_batch_tiling_entries(tmp_path, monkeypatch):
    """Table entries become [rows_lo, rows_hi, bins_lo, bins_hi + 1, shape,
    cap, lmax] (bins classes are inclusive); a forced shape disables them."""
    from riptide_amd import tune
    table = {"batch_tiles": [
        {"rows": [1, 128], "bins": [240, 260], "shape": 1, "cap": 0, "lmax": 3},
        {"rows": [128, 1 << 30], "bins": [480, 520], "shape": 2, "lmax": 4}]}
    path = tmp_path / "tuning.json"
    path.write_text(json.dumps(table))
    monkeypatch.setenv("RIPTIDE_AMD_TUNING", str(path))
    tune._cache = None
    try:
        assert tune.batch_tiling() == [[1, 128, 240, 261, 1, 0, 3],
                                       [128, 1 << 30, 480, 521, 2, 0, 4]]
        monkeypatch.setattr(tune, "_forced", 0)
        assert tune.batch_tiling() == []
    finally:
        tune._cache = None


def test_rows_period_range():
    """rows ~ nsamp * tsamp / period: the period range of a rows class, cut
    to the search range, or None when the class lies outside it."""
    from riptide_amd import tune
    span = 2 ** 22 * 64e-6
    assert tune.rows_period_range((1, 128), 2 ** 22, 64e-6, 0.1, 10.0) == (span / 128, 10.0)
    assert tune.rows_period_range((128, 1024), 2 ** 22, 64e-6, 0.1, 10.0) == (span / 1024,
                                                                                span / 128)
    assert tune.rows_period_range((1024, 1 << 30), 2 ** 22, 64e-6, 0.1, 10.0) == (0.1,
                                                                                  span / 1024)
    assert tune.rows_period_range((1, 128), 2 ** 22, 64e-6, 0.1, 1.0) is None


def test_batch_tiling_table_checked():
    """rt_batch_set_tiling rejects malformed entries with the library's
    error, and an empty table clears it."""
    from riptide_amd import _lib, tune
    try:
        _lib.lib()
    except Exception:
        pytest.skip("native library not built")
    with pytest.raises(ValueError, match="bad tiling table entry"):
        tune.set_batch_tiling([[0, 128, 240, 261, 5, 0, 3]])   # no shape 5
    with pytest.raises(ValueError, match="bad tiling table entry"):
        tune.set_batch_tiling([[128, 1, 240, 261, 0, 0, 3]])   # empty rows range
    with pytest.raises(ValueError, match="bad tiling table entry"):
        tune.set_batch_tiling([[0, 128, 240, 261, 0, 0, 9]])   # lmax beyond the planner
    tune.set_batch_tiling([[0, 128, 240, 261, 1, 0, 3]])
    tune.set_batch_tiling([])
