"""bench.py's host-side contract, on the CPU: the one-GPU exit and the HBM traffic reduction of
two rocprofv3 --pmc counter files (each counter scaled by what it reads for the 1 GiB copy
probe's dispatch)."""
import csv
import importlib.util
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_several_gpus_exit_with_a_message():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2
    line = json.loads(out.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def _counter_csv(path, counter, rows):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for k, v in rows:
            w.writerow({"Kernel_Name": k, "Counter_Name": counter, "Counter_Value": v})


def test_traffic_scaled_by_the_copy_probe(tmp_path):
    b = _bench()
    gib_kib = b.COPY_PROBE_BYTES / 1024.0
    z, y, x = "void scarl::k_h_strided<float, true, F>", "void scarl::k_h_strided<float, false, F>", \
        "void scarl::k_h_xpass<float, F, false, false>"
    # the copy kernel's reads register as half their bytes, its writes in full
    _counter_csv(tmp_path / "f.csv", "FETCH_SIZE",
                 [("__amd_rocclr_copyBuffer", gib_kib / 2), ("__amd_rocclr_copyBuffer", 3.0),
                  (z, 10.0), (y, 1000.0), (y, 1002.0), (x, 500.0)])
    _counter_csv(tmp_path / "w.csv", "WRITE_SIZE",
                 [("__amd_rocclr_copyBuffer", gib_kib), (z, 2000.0), (y, 2000.0), (x, 4000.0)])
    out = tmp_path / "traffic.json"
    b.traffic_from_counters(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), str(out))
    t = json.loads(out.read_text())
    assert t["fetch_factor"] == pytest.approx(2.0) and t["write_factor"] == pytest.approx(1.0)
    assert t["h_ypass"]["read_bytes"] == pytest.approx(1001.0 * 1024 * 2.0)
    assert t["h_xpass"]["write_bytes"] == pytest.approx(4000.0 * 1024)
    assert t["h_zpass"]["bytes"] == pytest.approx(10.0 * 1024 * 2.0 + 2000.0 * 1024)


def test_traffic_needs_the_copy_probe(tmp_path):
    b = _bench()
    _counter_csv(tmp_path / "f.csv", "FETCH_SIZE", [("void scarl::k_h_xpass<float, F>", 1.0)])
    _counter_csv(tmp_path / "w.csv", "WRITE_SIZE", [("void scarl::k_h_xpass<float, F>", 1.0)])
    with pytest.raises(SystemExit):
        b.traffic_from_counters(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), str(tmp_path / "t.json"))


def test_mfma_flop_count():
    """bf16 MFMA flops per generation: P pencils of a length N (multiple of 16, at least 32) cost
    P N / 256 groups of 16 columns x 12 products x 16384 flop; the 256-point x pass of a 512-wide
    field runs both of its radix-16 stages on the matrix cores."""
    b = _bench()
    per_group = 12 * 16 * 16 * 32 * 2
    z = 257 * 512 * (512 // 16 // 16) * per_group
    x = 512 * 512 * (256 // 16 // 16) * per_group * 2
    assert b._mfma_flops([(512, 512, 512)]) == 2 * z + x
    # 24 along x: the half-length 12 has no radix-16 stage; y = 16 is below the 32-point minimum
    assert b._mfma_flops([(24, 16, 64)]) == 13 * 16 * 64 * 768
    assert b._mfma_flops([(512, 512, 512)], x_only=True) == x
