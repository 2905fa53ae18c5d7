"""The saturating offsets of the device scans (curvedspacesim_amd/csrc/scan_math.h): candidate counts whose sum
passes INT32_MAX must not wrap the int32 offsets negative, no candidate may be written outside its buffer, and every
source that loses candidates must be flagged. The header's functions are constexpr and shared with the kernels; this
test compiles them into a host program with g++."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "curvedspacesim_amd", "csrc")


def test_saturating_offsets(tmp_path):
    exe = str(tmp_path / "scan_math_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I" + CSRC, os.path.join(HERE, "cpp", "scan_math_check.cpp"),
                    "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
