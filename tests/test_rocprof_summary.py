"""The profile summariser (gdmr.rocprof_summary) on small synthetic rocprofv3 CSV files: kernel
names shortened to their template-ids, per-dispatch means of counters that several rows (counter
instances) contribute to, the fetch-size calibration of the traffic summary and the traffic record
the benchmark reads."""
import csv
import json
import os

import pytest

from gdmr import rocprof_summary as rs

UPDATE = ("void dmr::blk::(anonymous namespace)::k_gemm_nn_b<1, float, 1>(dmr::bf16 const*, "
          "long long, int, float const*)")
COPY = "void dmr::(anonymous namespace)::k_scale_cols<float>(float*, long long, long long)"


def _counters(path, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "seq_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writerows(rows)


def test_short_names():
    assert rs.short_name(UPDATE) == "k_gemm_nn_b<1, float, 1>"
    assert rs.short_name(COPY) == "k_scale_cols<float>"
    assert rs.short_name("dmr::(anonymous namespace)::k_lsq(double const*, int)") == "k_lsq"


def test_counter_means_add_instances_then_average_dispatches(tmp_path):
    d = str(tmp_path / "run")
    # Two dispatches of the update, each reported as two instance rows; one of the copy.
    _counters(d, [(1, UPDATE, "TA_BUSY_avr", 10), (1, UPDATE, "TA_BUSY_avr", 20),
                  (2, UPDATE, "TA_BUSY_avr", 40), (2, UPDATE, "TA_BUSY_avr", 50),
                  (3, COPY, "TA_BUSY_avr", 7), (3, COPY, "GRBM_GUI_ACTIVE", 99)])
    means, _ = rs.counter_means(d, "TA_BUSY_avr")
    assert means["k_gemm_nn_b<1, float, 1>"] == (2, 60.0)
    assert means["k_scale_cols<float>"] == (1, 7.0)


def test_counters_mode_combines_runs(tmp_path, capsys):
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    _counters(a, [(1, UPDATE, "TA_BUSY_avr", 3), (2, UPDATE, "TA_BUSY_avr", 5)])
    _counters(b, [(1, UPDATE, "GRBM_GUI_ACTIVE", 80), (2, UPDATE, "GRBM_GUI_ACTIVE", 120)])
    rs.main(["counters", a, b, "--counters", "TA_BUSY_avr,GRBM_GUI_ACTIVE",
             "--kernels", "k_gemm_nn_b<1, float, 1>"])
    out = capsys.readouterr().out
    assert "| `k_gemm_nn_b<1, float, 1>` | 2 | 4 | 100 |" in out
    with pytest.raises(SystemExit):
        rs.main(["counters", a, "--counters", "NO_SUCH_COUNTER"])


def test_pmc_mode_calibrates_fetch_and_records_traffic(tmp_path, capsys):
    fd, wd = str(tmp_path / "fetch"), str(tmp_path / "write")
    # The copy kernel reads what it writes: WRITE 1000 KiB, FETCH 500 KiB -> factor 2.
    _counters(fd, [(1, COPY, "FETCH_SIZE", 500), (2, UPDATE, "FETCH_SIZE", 300)])
    _counters(wd, [(1, COPY, "WRITE_SIZE", 1000), (2, UPDATE, "WRITE_SIZE", 100)])
    rec = str(tmp_path / "traffic.json")
    rs.main(["pmc", fd, wd, "--calib", "k_scale_cols<float>", "--record",
             "k_gemm_nn_b<1, float, 1>", "--bytes", "700000", "--n", "8", "--s", "2",
             "--json", rec, "--source", "synthetic"])
    out = capsys.readouterr().out
    assert "fetch factor is x2.0000" in out
    r = json.load(open(rec))["k_gemm_nn_b<1, float, 1>"]
    assert r["traffic_bytes_per_launch"] == pytest.approx((300 * 2 + 100) * 1024.0)
    assert (r["n"], r["s"], r["bytes_per_launch"], r["source"]) == (8, 2, 700000, "synthetic")
