"""The cone passes' link-range plan (tsr_amd/csrc/link_ranges.h) on the CPU, with record
totals above 2^32: ranges stay below the record cap and 32-bit offsets, cover every link once
in order, merge whenever there is more than one, and a single link over the cap is refused
(tests/cpp/link_ranges_check.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_link_range_plan(tmp_path):
    exe = str(tmp_path / "link_ranges_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(ROOT, "tests", "cpp", "link_ranges_check.cpp")],
                   check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout
