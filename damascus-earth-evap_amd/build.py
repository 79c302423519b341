"""Build libdmsun.so (gfx950) in-tree with hipcc, then the C++ host API over its
C ABI (libdmsun_host.so) and the dmsun_run driver with g++.

Usage: python3 build.py        (from any directory)

Flags:
  -ffp-contract=off   keeps every FP64 expression a separate IEEE operation on
                      host and device, so kernel results match the CPU
                      restatement bit for bit;
  -packed-fp32-ops    disabled (no v_pk_* FP32 forms are emitted).
"""
import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libdmsun.so"
HOST = HERE / "host"
HOST_LIB = HERE / "libdmsun_host.so"
RUNNER = HERE / "dmsun_run"
SERVER = HERE / "dmsun_serve"
CXX = os.environ.get("CXX", "g++")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else "hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wno-unused-result", "-Wno-unused-value",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False):
    cmd = [HIPCC, f"--offload-arch={ARCH}", *COMMON, "-shared", "-o", str(LIB) + ".tmp",
           str(CSRC / "dms_runtime.hip"), str(CSRC / "dms_tables.cpp")]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libdmsun.so")
    os.replace(str(LIB) + ".tmp", LIB)
    build_host(verbose)
    return LIB


def build_host(verbose=False):
    """C++ host API (Solar_Model, Trajectory_Simulator, Simulation_Data,
    Configuration) and the command-line driver, linked against libdmsun.so."""
    link = ["-L", str(HERE), "-ldmsun", "-Wl,-rpath,$ORIGIN"]
    steps = [
        [CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-o", str(HOST_LIB) + ".tmp", str(HOST / "dmsun_output.cpp"),
         str(HOST / "dmsun_host.cpp"), str(HOST / "dmsun_server.cpp"), "-pthread", *link],
        [CXX, "-O2", "-std=c++17", "-o", str(RUNNER) + ".tmp", str(HOST / "dmsun_main.cpp"), "-L", str(HERE),
         "-ldmsun_host", "-ldmsun", "-Wl,-rpath,$ORIGIN"],
        [CXX, "-O2", "-std=c++17", "-pthread", "-o", str(SERVER) + ".tmp", str(HOST / "dmsun_serve.cpp"), "-L", str(HERE),
         "-ldmsun_host", "-ldmsun", "-Wl,-rpath,$ORIGIN"],
    ]
    for cmd, out in zip(steps, (HOST_LIB, RUNNER, SERVER)):
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"{CXX} failed building {out.name}")
        os.replace(str(out) + ".tmp", out)
    return HOST_LIB


if __name__ == "__main__":
    print(build(verbose=True))
