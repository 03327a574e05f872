"""The rocprofv3 summary tool on the CPU: kernel lists keep commas inside template
arguments, and the traffic summary scales FETCH_SIZE / WRITE_SIZE by the copy
calibration."""
import csv
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "multigraph_amd"))

import profile_summary  # noqa: E402


def test_split_names():
    assert profile_summary.split_names("k_a<unsigned char, unsigned short>,k_b,k_c<int>") == [
        "k_a<unsigned char, unsigned short>", "k_b", "k_c<int>"]
    assert profile_summary.split_names("k_x") == ["k_x"]


def _pmc(dirname, counter, rows):
    os.makedirs(dirname)
    with open(os.path.join(dirname, "x_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"])
        for i, (name, value) in enumerate(rows):
            w.writerow([i + 1, name, 256, counter, value])


def test_traffic_scaling(tmp_path):
    mib = 1024 * 1024
    # copy kernel: 1024 MiB read and written; FETCH_SIZE reports half the reads (KiB)
    _pmc(tmp_path / "cf", "FETCH_SIZE", [("k_copy(float*)", 512 * 1024)] * 3)
    _pmc(tmp_path / "cw", "WRITE_SIZE", [("k_copy(float*)", 1024 * 1024)] * 3)
    name = "k_ps_gather<unsigned char, unsigned short>(int*)"
    _pmc(tmp_path / "f", "FETCH_SIZE", [(name, 100 * 1024)] * 2)
    _pmc(tmp_path / "w", "WRITE_SIZE", [(name, 50 * 1024)] * 2)
    out = subprocess.run([sys.executable, os.path.join(ROOT, "multigraph_amd", "profile_summary.py"), "traffic",
                          "--fetch", str(tmp_path / "f"), "--write", str(tmp_path / "w"),
                          "--copy-fetch", str(tmp_path / "cf"), "--copy-write", str(tmp_path / "cw"),
                          "--copy-mib", "1024", "--kernels", "k_ps_gather<unsigned char, unsigned short>"],
                         capture_output=True, text=True, check=True).stdout
    row = [l for l in out.splitlines() if l.startswith("| k_ps_gather")]
    assert row, out
    cells = [c.strip() for c in row[0].split("|")[1:-1]]
    assert cells[1] == "2"
    # FETCH_SIZE counts half the copy's reads: 100 MiB reported per launch = 200 MiB read
    assert abs(float(cells[2]) - 200 * mib / 1e6) < 0.2
    assert abs(float(cells[3]) - 50 * mib / 1e6) < 0.2
