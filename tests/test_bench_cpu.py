"""bench.py's helpers that read the committed rocprofv3 summaries (no GPU needed): the far
pass's counter-measured HBM traffic and VALU counters that the JSON line reports."""
import importlib.util
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_far_traffic_and_valu_come_from_the_profile():
    bench = _bench()
    traffic = bench.far_traffic()
    assert traffic is not None and 1e8 < traffic < 1e9           # bytes per far pass at N = 2000
    valu = bench.far_valu()
    assert valu is not None
    assert 0.0 < valu["valu_busy_percent"] <= 100.0
    assert 0.0 < valu["valu_utilization_percent"] <= 100.0
    assert valu["source"].startswith("profiles/")


def test_gpus_above_one_is_refused(capsys, monkeypatch):
    bench = _bench()
    monkeypatch.setattr("sys.argv", ["bench.py", "--gpus", "2"])
    try:
        bench.main()
    except SystemExit as e:
        assert e.code == 2
    out = capsys.readouterr().out
    assert "one GPU" in out
