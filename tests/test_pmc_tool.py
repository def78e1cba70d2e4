"""The rocprofv3 summary tool (mbsymmetry_amd/pmc.py) on small CSV files in
rocprofv3's layout: kernel statistics, per-dispatch counter means and the
calibrated traffic figure."""
import csv
import json
import os

from mbsymmetry_amd import pmc

COUNTER_FIELDS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
                  "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
                  "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp",
                  "End_Timestamp"]


def _counter_csv(path, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=COUNTER_FIELDS)
        w.writeheader()
        for disp, name, counter, value in rows:
            r = {k: "0" for k in COUNTER_FIELDS}
            r.update(Dispatch_Id=str(disp), Correlation_Id=str(disp), Kernel_Name=name, Counter_Name=counter,
                     Counter_Value=str(value))
            w.writerow(r)


SIG = "void symgw::sigma_mfma_kernel<double, 2, false>(symgw::SigItem const*, long)"
MUL = "void at::native::vectorized_elementwise_kernel<4, at::native::MulFunctor<double> >(int, double*)"
FILL = "void at::native::vectorized_elementwise_kernel<4, at::native::FillFunctor<double> >(int, double*)"


def test_stats_table(tmp_path):
    d = tmp_path / "trace" / "host"
    d.mkdir(parents=True)
    with open(d / "p_kernel_stats.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs", "StdDev"])
        w.writerow([SIG, 2, 7.8e9, 3.9e9, 61.5, 3.8e9, 4.0e9, 1e7])
        w.writerow(["symgw::wsolve_kernel(int)", 2, 1.0e8, 5e7, 0.8, 4e7, 6e7, 1e6])
    rows = pmc.kernel_stats(str(tmp_path / "trace"))
    assert rows[0]["kernel"] == "void symgw::sigma_mfma_kernel<double, 2, false>"
    assert rows[0]["calls"] == 2 and abs(rows[0]["avg_ms"] - 3900.0) < 1e-9


def test_counter_means_sum_rows_of_a_dispatch(tmp_path):
    # two dispatches; the second written as two partial rows (summed)
    _counter_csv(str(tmp_path / "c" / "p_counter_collection.csv"),
                 [(1, SIG, "FETCH_SIZE", 100.0), (2, SIG, "FETCH_SIZE", 150.0), (2, SIG, "FETCH_SIZE", 50.0),
                  (3, FILL, "FETCH_SIZE", 7.0)])
    res = pmc.counters([str(tmp_path / "c")], "sigma")
    assert list(res) == ["void symgw::sigma_mfma_kernel<double, 2, false>"]
    v = res["void symgw::sigma_mfma_kernel<double, 2, false>"]
    assert v["FETCH_SIZE"] == 150.0 and v["dispatches"] == 2


def test_traffic_is_calibrated_by_the_copy(tmp_path):
    gib = 1 << 30
    # the copy: two dispatches of 2 GiB read and 2 GiB written; FETCH_SIZE reads half the bytes
    _counter_csv(str(tmp_path / "cf" / "a_counter_collection.csv"),
                 [(1, FILL, "FETCH_SIZE", 0.0), (2, MUL, "FETCH_SIZE", gib / 1024), (3, MUL, "FETCH_SIZE", gib / 1024)])
    _counter_csv(str(tmp_path / "cw" / "a_counter_collection.csv"),
                 [(1, FILL, "WRITE_SIZE", 4 * gib / 1024), (2, MUL, "WRITE_SIZE", 2 * gib / 1024),
                  (3, MUL, "WRITE_SIZE", 2 * gib / 1024)])
    _counter_csv(str(tmp_path / "f" / "a_counter_collection.csv"), [(9, SIG, "FETCH_SIZE", 1000.0)])
    _counter_csv(str(tmp_path / "w" / "a_counter_collection.csv"), [(9, SIG, "WRITE_SIZE", 10.0)])
    t = pmc.traffic(str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "cf"), str(tmp_path / "cw"),
                    "sigma_mfma", calib_kernel="MulFunctor")
    assert abs(t["fetch_scale"] - 2.0) < 1e-12 and abs(t["write_scale"] - 1.0) < 1e-12
    assert abs(t["bytes_per_launch"] - (2.0 * 1000 * 1024 + 10 * 1024)) < 1e-6
    out = tmp_path / "t.json"
    assert pmc.main(["traffic", "--fetch", str(tmp_path / "f"), "--write", str(tmp_path / "w"), "--calib-fetch",
                     str(tmp_path / "cf"), "--calib-write", str(tmp_path / "cw"), "--kernel", "sigma_mfma",
                     "--calib-kernel", "MulFunctor", "--out", str(out)]) == 0
    assert json.load(open(out))["bytes_per_launch"] == t["bytes_per_launch"]


def test_short_names_keep_anonymous_namespaces():
    assert pmc._short("void symgw::(anonymous namespace)::mfma_peak_kernel<double>(double*, int)") == \
        "void symgw::(anonymous namespace)::mfma_peak_kernel<double>"
    assert pmc._short("symgw::tgemm_kernel(symgw::TBatch const*, int)") == "symgw::tgemm_kernel"
    assert pmc._short("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"
