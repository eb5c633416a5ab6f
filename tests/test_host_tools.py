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
        # a 512 MiB copy measured at 2.25x its bytes, a larger concatenation (not a
        # calibration candidate) and two chemistry dispatches
        ("__amd_rocclr_copyBuffer", "FETCH_SIZE", 1, 512 * 1024 * 2.25),
        ("void at::native::CatArrayBatchedCopy_contig<double>", "FETCH_SIZE", 4, 4096 * 1024.0),
        # smaller dispatches of the same copy kernel (field clones) do not dilute it
        ("void at::native::elementwise_kernel_manual_unroll<128, 4, at::native::direct_copy_kernel_cuda>",
         "FETCH_SIZE", 5, 1024.0),
        ("void at::native::elementwise_kernel_manual_unroll<128, 4, at::native::direct_copy_kernel_cuda>",
         "FETCH_SIZE", 6, 2048.0),
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
    f.write_text("# command: rocprofv3 ... -- python3 bench.py --chunk-log2 1\n"
                 "TRAFFIC\tchem_bdf_kernel\tFETCH_SIZE\t1000\n"
                 "TRAFFIC\tchem_bdf_kernel\tWRITE_SIZE\t500\nTRAFFIC\trates_kernel\tWRITE_SIZE\t100\n"
                 "TRAFFIC\tclosure_kernel\tWRITE_SIZE\t7\n")
    monkeypatch.setattr(bench, "HBM_PROFILE", str(f))
    t = bench.profiled_traffic(4)           # 4 cells, profiled dispatches of 2 cells
    assert t["bytes_per_call"] == 2 * 1600
    assert t["profiled_dispatch_cells"] == 2
    assert set(t["per_dispatch"]) == {"chem_bdf_kernel FETCH_SIZE", "chem_bdf_kernel WRITE_SIZE",
                                      "rates_kernel WRITE_SIZE"}


def test_set_gate_rewrites_only_the_thresholds(tmp_path):
    """train_surrogate --set-gate: the named gate thresholds change, every other
    array of the trained file (parameters, normalizer, metadata) is kept."""
    import numpy as np
    from nrg_amd import train_surrogate
    src = tmp_path / "in.npz"
    np.savez(src, params=np.arange(7, dtype=np.float32), in_shift=np.ones(12, np.float32),
             gate_dT_max=np.float64(0.25), gate_energy_tol=np.float64(0.1),
             meta=np.array(repr({"seed": 3})))
    out = tmp_path / "out"
    assert train_surrogate.main(["--out", str(out), "--set-gate", str(src),
                                 "--gate", "dT_max=10"]) == 0
    with np.load(src) as a, np.load(out / "surrogate_keromnes.npz") as b:
        assert sorted(a.files) == sorted(b.files)
        assert float(b["gate_dT_max"]) == 10.0
        for k in a.files:
            if k != "gate_dT_max":
                assert np.array_equal(a[k], b[k]), k
    import pytest
    with pytest.raises(KeyError):
        train_surrogate.main(["--out", str(out), "--set-gate", str(src), "--gate", "nope=1"])


def test_counter_summary_last_dispatches_and_lane_share(tmp_path):
    """--last N averages each kernel's last N dispatches only (the timed steps),
    and the lane share is SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)."""
    from nrg_amd import rocprof_summary as rs
    p = tmp_path / "sq.csv"
    k = "void (anonymous namespace)::chem_bdf_kernel<true, false>((anonymous namespace)::StepArgs)"
    rows = [(k, "SQ_ACTIVE_INST_VALU", 1, 1000.0), (k, "SQ_THREAD_CYCLES_VALU", 1, 64000.0),
            (k, "SQ_ACTIVE_INST_VALU", 7, 100.0), (k, "SQ_THREAD_CYCLES_VALU", 7, 1600.0),
            (k, "SQ_ACTIVE_INST_VALU", 9, 300.0), (k, "SQ_THREAD_CYCLES_VALU", 9, 4800.0)]
    with open(p, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Counter_Name", "Dispatch_Id", "Counter_Value"])
        for r in rows:
            w.writerow(r)
    buf = io.StringIO()
    with redirect_stdout(buf):
        rs.counters(str(p), last=2)
    out = buf.getvalue()
    lanes = [l.split("\t") for l in out.splitlines() if l.startswith("LANES")]
    # dispatches 7 and 9 only: (1600 + 4800) / (64 * 400) = 0.25
    assert lanes == [["LANES", "chem_bdf_kernel", "0.2500"]]
    avg = [l.split() for l in out.splitlines() if "SQ_ACTIVE_INST_VALU" in l]
    assert avg and float(avg[0][-1]) == 200.0 and avg[0][-3] == "2"
