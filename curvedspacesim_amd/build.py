"""Builds the native libraries in-tree: libcsg.so (HIP kernels + C-ABI, gfx950), libcsg_host.so (C++ host layer) and,
where the HDF5 C library is found, libcsg_h5.so (trajectory files in the reference's simpleModelDatabase layout)."""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
ROOT = os.path.dirname(PKG)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["csg_api.hip", "csg_system.hip", "csg_screen.hip", "geodesic_kernel.hip", "transport_kernel.hip", "screen_kernel.hip"]
HOST_SOURCES = ["host/meshGenerators.cpp", "host/hostCapi.cpp", "host/noiseSource.cpp", "host/gpuMeshSpace.cpp",
                "host/neighborStructures.cpp", "host/simpleModel.cpp", "host/forces.cpp", "host/simulation.cpp"]
# reference-style scenario driver of the C++ host classes, run by the parity tests
SCENARIOS = os.path.join(ROOT, "tests", "cpp", "refapi_scenarios.cpp")
# packed FP32 instructions are kept out of every code object (see the hipcc feature flag below); the compiler's
# wavefront-level combining of same-address atomics is off: the geodesic kernel appends work items from divergent
# branches with few lanes active, where the scan it inserts costs more than the atomics it saves (measured 2 %)
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-mllvm", "-amdgpu-atomic-optimizer-strategy=None"]


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_csg():
    out = os.path.join(PKG, "libcsg.so")
    cmd = [HIPCC] + HIP_FLAGS + [os.path.join(CSRC, s) for s in HIP_SOURCES] + ["-o", out + ".tmp"]
    _run(cmd, CSRC)
    os.replace(out + ".tmp", out)
    return out


def build_host():
    out = os.path.join(PKG, "libcsg_host.so")
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall",
           "-I" + os.path.join(ROOT, "include")] + [os.path.join(CSRC, s) for s in HOST_SOURCES] + \
          ["-L" + PKG, "-lcsg", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
    _run(cmd, CSRC)
    os.replace(out + ".tmp", out)
    return out


# the HDF5 C library: CSG_HDF5_DIR, else the image's conda prefix (HDF5 1.10)
HDF5_DIR = os.environ.get("CSG_HDF5_DIR", "/opt/conda")


def hdf5_found():
    return os.path.exists(os.path.join(HDF5_DIR, "include", "hdf5.h")) and os.path.exists(os.path.join(HDF5_DIR, "lib", "libhdf5.so"))


def build_h5():
    """libcsg_h5.so, or None (after removing a stale one) when the HDF5 C library is not found."""
    out = os.path.join(PKG, "libcsg_h5.so")
    if not hdf5_found():
        if os.path.exists(out):
            os.remove(out)
        print(f"HDF5 C library not found under {HDF5_DIR}: libcsg_h5.so (HDF5 trajectory files) not built")
        return None
    srcs = [os.path.join(CSRC, "host", f) for f in ("simpleModelDatabase.cpp", "h5Capi.cpp")]
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(HDF5_DIR, "include")] + srcs + \
          ["-L" + PKG, "-lcsg_host", "-lcsg", "-L" + os.path.join(HDF5_DIR, "lib"), "-lhdf5",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(HDF5_DIR, "lib"), "-o", out + ".tmp"]
    _run(cmd, CSRC)
    os.replace(out + ".tmp", out)
    return out


def build_scenarios():
    out = os.path.join(ROOT, "tests", "cpp", "refapi_scenarios")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(CSRC, "host"), SCENARIOS, "-L" + PKG, "-lcsg_host", "-lcsg",
           "-Wl,-rpath," + os.path.relpath(PKG, os.path.dirname(out)).join(["$ORIGIN/", ""]), "-o", out + ".tmp"]
    _run(cmd, CSRC)
    os.replace(out + ".tmp", out)
    return out


def build_all(verbose=False):
    a = build_csg()
    b = build_host()
    h = build_h5()
    c = build_scenarios()
    if verbose:
        print("built", a, b, c, h or "(no HDF5 library)")
    return a, b, c


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
