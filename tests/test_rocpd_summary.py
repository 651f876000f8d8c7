"""The profile summary tool: per-kernel means from rocprofv3 CSV output, with
each kernel's first dispatches left out on request."""

import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "gkmsvm-lite_amd"))

from gkmsvm_amd import rocpd_summary as R  # noqa: E402

TRACE = '''"Kind","Agent_Id","Queue_Id","Kernel_Id","Kernel_Name","Correlation_Id","Start_Timestamp","End_Timestamp"
"KERNEL_DISPATCH",1,1,1,"k_a(int)",2,300000,700000
"KERNEL_DISPATCH",1,1,1,"k_a(int)",1,100000,200000
"KERNEL_DISPATCH",1,1,2,"k_b",3,800000,1000000
"KERNEL_DISPATCH",1,1,2,"k_b",4,1100000,1400000
'''
PMC = '''"Correlation_Id","Dispatch_Id","Kernel_Name","Counter_Name","Counter_Value"
2,2,"k_a(int)","FETCH_SIZE",10
1,1,"k_a(int)","FETCH_SIZE",2
3,3,"k_b","FETCH_SIZE",7
'''


def _write(tmp_path):
    d = tmp_path / "host" / "123"
    d.mkdir(parents=True)
    (d / "run_kernel_trace.csv").write_text(TRACE)
    (d / "run_counter_collection.csv").write_text(PMC)
    return str(tmp_path)


def _rows(text):
    return {ln.split()[0]: ln.split()[1:] for ln in text.splitlines()[1:]}


def test_trace_csv_means_and_skips(tmp_path):
    d = _write(tmp_path)
    rows = _rows(R.trace_csv(d))
    assert rows["k_a"][:1] == ["2"] and float(rows["k_a"][2]) == 250.0  # (100 + 400) us / 2
    rows = _rows(R.trace_csv(d, skip=1))  # the first dispatch (by start time) of each
    assert rows["k_a"][0] == "1" and float(rows["k_a"][2]) == 400.0
    assert rows["k_b"][0] == "1" and float(rows["k_b"][2]) == 300.0
    rows = _rows(R.trace_csv(d, skip="k_a=1"))  # by kernel name
    assert rows["k_a"][0] == "1" and rows["k_b"][0] == "2"


def test_pmc_skips_in_dispatch_order(tmp_path):
    d = _write(tmp_path)
    assert float(_rows(R.pmc(d, "k_"))["k_a"][2]) == 6.0
    rows = _rows(R.pmc(d, "k_", "k_a=1"))
    assert float(rows["k_a"][2]) == 10.0 and float(rows["k_b"][2]) == 7.0


def test_traffic_scales_counters_by_the_copy_calibration(tmp_path):
    """HBM bytes = FETCH_SIZE x (copy bytes / copy FETCH_SIZE) + WRITE_SIZE x
    (copy bytes / copy WRITE_SIZE), template instances pooled by base name;
    bench.py reads the TRAFFIC / VALU lines back."""
    def run(name, rows):
        d = tmp_path / name / "h"
        d.mkdir(parents=True)
        body = '"Correlation_Id","Dispatch_Id","Kernel_Name","Counter_Name","Counter_Value"\n'
        body += "".join(f'{i},{i},"{k}","{c}",{v}\n' for i, (k, c, v) in enumerate(rows, 1))
        (d / "run_counter_collection.csv").write_text(body)
        return str(tmp_path / name)

    cb = R.COPY_BYTES // 1024
    cf = run("cf", [("__amd_rocclr_copyBuffer", "FETCH_SIZE", cb / 2)])
    cw = run("cw", [("__amd_rocclr_copyBuffer", "WRITE_SIZE", cb)])
    f = run("f", [("void k_x<4, 1>(A)", "FETCH_SIZE", 100), ("void k_x<2, 0>(A)", "FETCH_SIZE", 300),
                  ("k_y", "FETCH_SIZE", 10)])
    w = run("w", [("void k_x<4, 1>(A)", "WRITE_SIZE", 8), ("void k_x<2, 0>(A)", "WRITE_SIZE", 8),
                  ("k_y", "WRITE_SIZE", 0)])
    v = run("v", [("void k_x<4, 1>(A)", "SQ_INSTS_VALU", 640)])
    text = R.traffic(f, w, cf, cw, 0, v)
    assert "FETCH_SIZE x2.000 WRITE_SIZE x1.000" in text
    p = tmp_path / "summary.txt"
    p.write_text("header\n" + text + "\n")
    prof = R.read_profile(str(p))
    assert prof["traffic"]["k_x"] == (200 * 2 + 8) * 1024.0
    assert prof["traffic"]["k_y"] == 10 * 2 * 1024.0
    assert prof["valu"] == {"k_x": 640.0}
