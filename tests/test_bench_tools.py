"""The counter and kernel-trace summaries bench.py writes from rocprofv3's
CSV files: only the timed walks after the counting launch are averaged,
FETCH_SIZE / WRITE_SIZE are scaled by what they read for the 512 MiB
calibration copy, and the kernel table names every kernel (CPU only, on
synthetic CSV files)."""
import csv
import json
import shutil

import pytest

import bench

WALK = "void rt::(anonymous namespace)::k_trace<false, false, false>(rt::(anonymous namespace)::TraceArgs)"
COUNT = "void rt::(anonymous namespace)::k_trace<true, false, false>(rt::(anonymous namespace)::TraceArgs)"
COPY = "__amd_rocclr_copyBuffer"
HDR = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _counters(path, launches):
    """launches: [(kernel, {counter: value})] in dispatch order."""
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(HDR)
        for i, (k, vals) in enumerate(launches, 1):
            for c, v in vals.items():
                w.writerow([i, k, c, v])


@pytest.fixture
def profile_dir(tmp_path, monkeypatch):
    """A repository root with an empty profiles/ and the CSV files of one
    ray-cast profiling session: a foreign walk before the counting launch
    (ignored), the counting launch, the copy, then two timed walks."""
    root = tmp_path / "root"
    (root / "profiles").mkdir(parents=True)
    for f in bench.KERNEL_SOURCES:
        (root / f).parent.mkdir(parents=True, exist_ok=True)
        shutil.copy(bench.ROOT / f, root / f)
    monkeypatch.setattr(bench, "ROOT", root)
    d = tmp_path / "pmc"
    d.mkdir()
    sq = ("SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CYCLES", "SQ_WAVES", "SQ_WAVE_CYCLES",
          "SQ_INSTS_VMEM_RD", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "TCC_EA0_RDREQ_sum")

    def session(name, timed):
        _counters(d / f"raycast_{name}_counter_collection.csv",
                  [(WALK, {c: 9e9 for c in timed[0]}), (COUNT, {c: 1.0 for c in timed[0]}),
                   (COPY, timed[1])] + [(WALK, t) for t in timed[2:]])

    session("fetch", [{"FETCH_SIZE": 0}, {"FETCH_SIZE": 262144.0}, {"FETCH_SIZE": 1000.0}, {"FETCH_SIZE": 3000.0}])
    session("write", [{"WRITE_SIZE": 0, "TCC_HIT_sum": 0, "TCC_MISS_sum": 0}, {"WRITE_SIZE": 524288.0},
                      {"WRITE_SIZE": 100.0, "TCC_HIT_sum": 30.0, "TCC_MISS_sum": 10.0},
                      {"WRITE_SIZE": 300.0, "TCC_HIT_sum": 30.0, "TCC_MISS_sum": 10.0}])
    sqv = dict({c: 1000.0 for c in sq}, GRBM_GUI_ACTIVE=8 * 500.0)
    session("sq", [sqv, {"SQ_INSTS_VALU": 5.0}, dict(sqv, SQ_INSTS_VALU=64000.0), dict(sqv, SQ_INSTS_VALU=64000.0)])
    with open(d / "raycast_stats_kernel_stats.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs", "StdDev"])
        w.writerow([WALK, 3, 15e6, 5e6, 90, 0, 0, 0])
        w.writerow([COUNT, 1, 2e6, 2e6, 10, 0, 0, 0])
    with open(d / "raycast_stats_kernel_trace.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writerow([1, WALK, 0, 9_000_000])       # before the counting launch: not a timed walk
        w.writerow([2, COUNT, 0, 2_000_000])
        w.writerow([3, WALK, 0, 4_000_000])
        w.writerow([4, WALK, 0, 6_000_000])
    return root, d


def test_pmc_summary_uses_the_timed_walks_and_the_copy_calibration(profile_dir):
    root, d = profile_dir
    bench.summarize_pmc("raycast", d)
    out = json.loads((root / "profiles" / "trace_raycast_1M_16M_pmc.json").read_text())
    assert out["launches_counted"] == 2 and out["kernel_launches_traced"] == 2
    assert out["fetch_factor"] == 2.0 and out["write_factor"] == 1.0   # 512 MiB / (KiB read by the copy)
    assert out["fetch_kib_raw"] == 2000.0 and out["write_kib_raw"] == 200.0
    assert out["hbm_bytes_per_launch"] == (2000 * 2 + 200) * 1024
    assert out["l2_hit_rate"] == 0.75
    assert out["kernel_ms_trace"] == 5.0
    # 64000 wave instructions over 256 CUs x 500 cycles per XCD
    assert out["valu_insts_per_cu_cycle"] == round(64000 / (256 * 500), 4)
    assert out["valu_busy"] == round(1000 / (256 * 500), 4)
    assert out["source_hash"] == bench.kernel_source_hash()
    md = (root / "profiles" / "trace_raycast_1M_16M_kernel_stats.md").read_text()
    assert "`k_trace<false, false, false>` | 3 |" in md and "mean 5.0000 ms" in md


def test_kernel_stats_table_names_every_kernel(profile_dir):
    root, d = profile_dir
    bench.summarize_stats(d, "walks", "rocprofv3 --kernel-trace --stats -- python3 bench.py")
    md = (root / "profiles" / "walks_kernel_stats.md").read_text()
    rows = [r for r in md.splitlines() if r.startswith("| `")]
    assert [r.split("`")[1] for r in rows] == ["k_trace<false, false, false>", "k_trace<true, false, false>"]
    assert "rocprofv3 --kernel-trace --stats -- python3 bench.py" in md
