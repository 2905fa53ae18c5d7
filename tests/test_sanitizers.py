"""The oracle and the CPU-only pieces of the host layer (mesh generators and OFF I/O, noise source, cell list) built
with AddressSanitizer and UndefinedBehaviorSanitizer and driven through their main and error paths by
tests/cpp/sanitize_driver.cpp; any memory error, leak or undefined behaviour fails the run."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE = os.path.join(ROOT, "oracle")
HOST = os.path.join(ROOT, "curvedspacesim_amd", "csrc", "host")


@pytest.mark.slow
def test_oracle_and_host_under_sanitizers(tmp_path):
    exe = str(tmp_path / "sanitize_driver")
    srcs = [os.path.join(ORACLE, f) for f in ("mesh.cpp", "geodesic.cpp", "walk.cpp", "model.cpp")]
    srcs += [os.path.join(HOST, f) for f in ("meshGenerators.cpp", "noiseSource.cpp", "neighborStructures.cpp")]
    srcs.append(os.path.join(HERE, "cpp", "sanitize_driver.cpp"))
    flags = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fopenmp", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=all"]
    subprocess.run(["g++"] + flags + ["-I" + ORACLE, "-I" + HOST] + srcs + ["-o", exe], check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), (r.stdout[-3000:] + r.stderr[-3000:])
