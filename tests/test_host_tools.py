"""Host-side measurement plumbing: the rocprof summary's calibrated TRAFFIC
lines and the bench's per-call HBM traffic built from them."""
import csv
import importlib
import io
import os
import sys
from contextlib import redirect_stdout

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_counter_summary_calibrates_and_emits_traffic(tmp_path):
    from nrg_amd import rocprof_summary as rs
    p = tmp_path / "c.csv"
    rows = [
        # a 512 MiB copy measured at 2.25x its bytes, and two chemistry dispatches
        ("void at::native::elementwise_kernel<copy>", "FETCH_SIZE", 1, 512 * 1024 * 2.25),
        ("void (anonymous namespace)::chem_bdf_kernel<true, false>((anonymous namespace)::StepArgs)",
         "FETCH_SIZE", 2, 900.0),
        ("void (anonymous namespace)::chem_bdf_kernel<true, false>((anonymous namespace)::StepArgs)",
         "FETCH_SIZE", 3, 1350.0),
    ]
    with open(p, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Counter_Name", "Dispatch_Id", "Counter_Value"])
        for r in rows:
            w.writerow(r)
    buf = io.StringIO()
    with redirect_stdout(buf):
        rs.counters(str(p), calib_mib=512)
    out = buf.getvalue()
    assert "factor 0.4444" in out
    tr = [l.split("\t") for l in out.splitlines() if l.startswith("TRAFFIC")]
    chem = [t for t in tr if t[1] == "chem_bdf_kernel"][0]
    # (900 + 1350) / 2 dispatches / 2.25 KiB -> bytes
    assert abs(float(chem[3]) - 1125.0 / 2.25 * 1024) < 1.0


def test_bench_reads_profiled_traffic(tmp_path, monkeypatch):
    sys.path.insert(0, ROOT)
    bench = importlib.import_module("bench")
    f = tmp_path / "hbm.txt"
    f.write_text("# summary\nTRAFFIC\tchem_bdf_kernel\tFETCH_SIZE\t1000\n"
                 "TRAFFIC\tchem_bdf_kernel\tWRITE_SIZE\t500\nTRAFFIC\trates_kernel\tWRITE_SIZE\t100\n"
                 "TRAFFIC\tclosure_kernel\tWRITE_SIZE\t7\n")
    monkeypatch.setattr(bench, "HBM_PROFILE", str(f))
    t = bench.profiled_traffic(2)
    assert t["bytes_per_call"] == 2 * 1600
    assert set(t["per_dispatch"]) == {"chem_bdf_kernel FETCH_SIZE", "chem_bdf_kernel WRITE_SIZE",
                                      "rates_kernel WRITE_SIZE"}
