"""CPU checks of the measurement helpers in bench.py: the algorithmic byte
model, the calibration of the HBM byte counters and the committed traffic
summary that the benchmark line quotes."""
import csv
import json
import os

import pytest

import bench


def _write_counters(path, counter, rows):
    """rows: (dispatch id, kernel name, value in KiB) as rocprofv3 writes them."""
    fields = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id",
              "Grid_Size", "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size",
              "Scratch_Size", "VGPR_Count", "Accum_VGPR_Count", "SGPR_Count", "Counter_Name",
              "Counter_Value", "Start_Timestamp", "End_Timestamp"]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for did, name, kib in rows:
            r = {k: 0 for k in fields}
            r.update(Dispatch_Id=did, Kernel_Name=name, Counter_Name=counter, Counter_Value=kib)
            w.writerow(r)


def test_canonical_bytes_128cubed_p4():
    # E element nodes, N global nodes of the 128^3 porder-4 box
    E, N = 128 ** 3 * 125, 513 ** 3
    bc, bd = bench.canonical_bytes(E, N, 8, 8)
    assert bc == E * 84 + N * 112
    assert bd == E * 100 + N * 88
    assert abs(bc / 1e9 - 37.14) < 0.01 and abs(bd / 1e9 - 38.09) < 0.01
    # FP32 geometry only changes the element terms
    bc32, bd32 = bench.canonical_bytes(E, N, 4, 8, 4)
    assert bc - bc32 == E * 40 and bd - bd32 == E * 48


def test_summarize_pmc_scales_by_the_copy(tmp_path):
    cb = bench.CALIB_BYTES
    conv = ("void msod::k_convec<4, msod::ConvecArgs<double, double, double, double, double>, double>"
            "(msod::ConvecArgs<double, double, double, double, double>)")
    diff = ("void msod::k_diffu<4, msod::DiffuArgs<double, double, double, double>, false, double>"
            "(msod::DiffuArgs<double, double, double, double>, float const*)")
    # the fetch counter reads half of what the copy moves; the write counter all of it
    fetch = tmp_path / "fetch.csv"
    write = tmp_path / "write.csv"
    _write_counters(fetch, "FETCH_SIZE", [(1, "__amd_rocclr_copyBuffer", cb / 2 / 1024),
                                          (2, "__amd_rocclr_copyBuffer", 3.0),
                                          (3, conv, 1000.0), (4, conv, 3000.0), (5, diff, 500.0)])
    _write_counters(write, "WRITE_SIZE", [(1, "__amd_rocclr_copyBuffer", cb / 1024),
                                          (3, conv, 10.0), (4, conv, 30.0), (5, diff, 7.0)])
    out = bench.summarize_pmc(str(fetch), str(write), 128, 4)
    assert out["calibration"]["FETCH_SIZE"]["factor"] == pytest.approx(0.5)
    assert out["calibration"]["WRITE_SIZE"]["factor"] == pytest.approx(1.0)
    c = out["kernels"]["full_convec_ijk"]
    assert c["dispatches"] == 2
    assert c["fetch_size_bytes"] == pytest.approx(2000.0 * 1024 / 0.5)
    assert c["write_size_bytes"] == pytest.approx(20.0 * 1024)
    assert c["traffic_bytes"] == pytest.approx(4000.0 * 1024 + 20.0 * 1024)
    d = out["kernels"]["full_diffusion_ijk"]
    assert d["traffic_bytes"] == pytest.approx(1000.0 * 1024 + 7.0 * 1024)
    assert d["algorithmic_bytes"] == bench.canonical_bytes(128 ** 3 * 125, 513 ** 3, 8, 8)[1]


def test_summarize_pmc_refuses_a_run_without_the_copy(tmp_path):
    fetch = tmp_path / "fetch.csv"
    write = tmp_path / "write.csv"
    _write_counters(fetch, "FETCH_SIZE", [(1, "__amd_rocclr_copyBuffer", 3.0)])
    _write_counters(write, "WRITE_SIZE", [(1, "__amd_rocclr_copyBuffer", 3.0)])
    with pytest.raises(SystemExit):
        bench.summarize_pmc(str(fetch), str(write), 128, 4)


def test_committed_traffic_summary_is_read_for_its_workload_only():
    assert os.path.exists(bench.TRAFFIC_FILE)
    lines = [l for l in open(bench.TRAFFIC_FILE) if l.startswith(bench.TRAFFIC_TAG)]
    assert len(lines) == 1
    rec = json.loads(lines[0][len(bench.TRAFFIC_TAG):])
    t = bench.load_traffic(128, 4, "full_diffusion_ijk")
    assert t == rec["kernels"]["full_diffusion_ijk"]["traffic_bytes"]
    # calibrated traffic is at least the bytes each kernel must move
    for k in ("full_convec_ijk", "full_diffusion_ijk"):
        assert rec["kernels"][k]["traffic_bytes"] >= rec["kernels"][k]["algorithmic_bytes"]
    assert bench.load_traffic(64, 4, "full_diffusion_ijk") is None
    assert bench.load_traffic(128, 3, "full_convec_ijk") is None


def test_minimal_flops_128cubed_p4():
    # 625 GFLOP per convection + diffusion pair at 128^3 porder 4
    fc, fd = bench.minimal_flops(128 ** 3 * 125, 5)
    assert abs((fc + fd) / 1e9 - 625.48) < 0.01
    blk = bench.flops_block(128 ** 3 * 125, 5, {"convec_ms": 10.0, "diffu_ms": 10.0})
    assert blk["convec_tflops"] == pytest.approx(fc / 1e-2 / 1e12)
    assert blk["pair_fp64_peak_frac"] == pytest.approx((fc + fd) / 2e-2 / bench.FP64_PEAK)


def test_summarize_counters_per_kernel(tmp_path):
    conv = ("void msod::k_convec<4, msod::ConvecArgs<double, double, double, double, double>, double>"
            "(msod::ConvecArgs<double, double, double, double, double>)")
    split = ("void msod::k_convec<4, msod::ConvecArgs<float, float, double, double, double>, double>"
             "(msod::ConvecArgs<float, float, double, double, double>)")
    a = tmp_path / "a.csv"
    b = tmp_path / "b.csv"
    _write_counters(a, "SQ_WAIT_ANY", [(1, conv, 30.0), (2, conv, 50.0), (3, split, 10.0)])
    _write_counters(b, "SQ_WAVE_CYCLES", [(1, conv, 100.0), (2, conv, 100.0), (3, split, 40.0)])
    table = bench.summarize_counters([str(a), str(b)]).splitlines()
    rows = {r.split("|")[1].strip(): [x.strip() for x in r.split("|")[2:-1]] for r in table[2:]}
    head = [x.strip() for x in table[0].split("|")[2:-1]]
    c = dict(zip(head, rows["k_convec<4, Convec<ddddd>, double>"]))
    assert c["dispatches"] == "2" and float(c["SQ_WAIT_ANY"]) == 40.0 and float(c["wait/wave"]) == 0.4
    f = dict(zip(head, rows["k_convec<4, Convec<ffddd>, double>"]))
    assert float(f["wait/wave"]) == 0.25
