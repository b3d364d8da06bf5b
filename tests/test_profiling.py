"""rocprofv3 summaries (CPU): the trace table keeps the dispatches after the
last dispatch of the named kernel and divides by the per-step kernel's
count."""
import csv

from voxvizthesis_amd import profiling


def _trace(path, names, dur_ns=1000):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        t = 0
        for n, d in names:
            w.writerow([f"voxmi::(anonymous namespace)::{n}(int, float*)", t, t + d])
            t += d + 500


def test_trace_after_last_and_per_step(tmp_path):
    p = tmp_path / "kt_kernel_trace.csv"
    _trace(p, [("k_a", 9000), ("k_copy_u128", 4000), ("k_b", 7000), ("k_copy_u128", 4000),
               ("k_a", 1000), ("k_step", 3000), ("k_a", 2000), ("k_step", 5000)])
    rows = profiling.trace_stats(str(p), "k_copy_u128", "k_step").splitlines()[2:]
    cells = {}
    for r in rows:
        c = [x.strip() for x in r.strip("|").split("|")]
        cells[c[0]] = c[1:]
    # before the last copy: k_b and the first k_a are left out
    assert "k_b" not in cells
    assert cells["k_a"][0] == "2" and cells["k_a"][1] == "1.5" and cells["k_a"][2] == "1.5"
    assert cells["k_step"][0] == "2" and cells["k_step"][2] == "4.0"
    assert cells["total (2 steps)"][2] == "5.5"


def test_trace_whole_run_without_options(tmp_path):
    p = tmp_path / "kt_kernel_trace.csv"
    _trace(p, [("k_a", 1000), ("k_b", 3000)])
    rows = profiling.trace_stats(str(p)).splitlines()[2:]
    assert [r.split(" | ")[0].lstrip("| ") for r in rows] == ["k_b", "k_a"]
    assert rows[0].endswith("| 75.0 |")
