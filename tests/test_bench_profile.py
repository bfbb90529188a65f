"""The benchmark's rocprofv3 summary: per (kernel, grid) statistics from
kernel-trace CSVs, and the call-weighted average that the roofline reads back
from a committed summary (selected by name prefix and grid size)."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def write_trace(path, rows):
    cols = ["Kind", "Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"]
    with open(path, "w") as fh:
        fh.write(",".join(f'"{c}"' for c in cols) + "\n")
        for name, t0, t1, gx in rows:
            fh.write(f'"KERNEL_DISPATCH","{name}",{t0},{t1},{gx},1,1\n')


def test_summary_and_weighted_average(bench, tmp_path, capsys, monkeypatch):
    d = tmp_path / "prof" / "host"
    d.mkdir(parents=True)
    write_trace(d / "run_kernel_trace.csv", [
        ("void moons::k_mg_line<1, 128, 512>(moons::LevelDev, int)", 0, 7000, 262144),
        ("void moons::k_mg_line<1, 128, 512>(moons::LevelDev, int)", 10000, 19000, 262144),   # 7 and 9 us -> avg 8
        ("k_mg_line<2, 128, 512>", 20000, 31000, 262144),   # 11 us
        ("k_mg_line<1, 64, 128>", 40000, 44000, 32768),     # another level: excluded
        ("k_pcg_update<8, false>", 50000, 70000, 262144),   # 20 us at 128^3
        ("k_pcg_update<8, false>", 80000, 81000, 4096),     # 1 us at 32^3: excluded by grid
    ])
    bench.rocprof_summary(str(tmp_path / "prof"))
    out = capsys.readouterr().out
    prof = tmp_path / "profiles"
    prof.mkdir()
    (prof / "s.txt").write_text("command: rocprofv3 ...\n" + out)
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    assert "kernel: void moons::k_mg_line<1, 128, 512>(moons::LevelDev, int) grid: 262144 calls: 2 avg_us: 8.000" in out
    # (2 * 8 + 1 * 11) / 3
    assert bench.load_rocprof_avg("s.txt", ("k_mg_line<1, 128,", "k_mg_line<2, 128,")) == pytest.approx(9.0)
    assert bench.load_rocprof_avg("s.txt", "k_pcg_update", grid=262144) == pytest.approx(20.0)
    assert bench.load_rocprof_avg("s.txt", "k_pcg_update") == pytest.approx(10.5)
    assert bench.load_rocprof_avg("s.txt", "k_absent") is None
    assert bench.load_rocprof_avg("missing.txt", "k_mg_line") is None


def test_counter_summary(bench, tmp_path, capsys):
    d = tmp_path / "pmc"
    d.mkdir()
    cols = ["Dispatch_Id", "Grid_Size", "Kernel_Name", "Counter_Name", "Counter_Value"]
    rows = [(1, 131072, "k_mg_line<1, 128, 512, 8>", "FETCH_SIZE", 4380.0),
            (2, 131072, "k_mg_line<1, 128, 512, 8>", "FETCH_SIZE", 4390.0),
            (3, 32768, "k_mg_line<1, 64, 128, 4>", "FETCH_SIZE", 614.0)]
    with open(d / "f_counter_collection.csv", "w") as fh:
        fh.write(",".join(f'"{c}"' for c in cols) + "\n")
        for r in rows:
            fh.write(f'{r[0]},{r[1]},"{r[2]}","{r[3]}",{r[4]}\n')
    bench.rocprof_summary(str(d))
    out = capsys.readouterr().out
    assert ("kernel: k_mg_line<1, 128, 512, 8> grid: 131072 counter: FETCH_SIZE dispatches: 2 "
            "mean: 4385.0 min: 4380.0 max: 4390.0") in out
    assert "grid: 32768 counter: FETCH_SIZE dispatches: 1 mean: 614.0" in out
