"""Builds tests/cpp/host_check (g++, host code only) against the C ABI library
and the oracle; the build is also run by __graft_entry__.build()."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "cpp", "host_check.cpp")
OUT_DIR = os.path.join(HERE, "cpp", "_build")
EXE = os.path.join(OUT_DIR, "host_check")
CSRC = os.path.join(ROOT, "02_cathode-08_modulerev_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")


def build_host_check(force=False):
    deps = [SRC] + [os.path.join(ROOT, "02_cathode-08_modulerev_amd", "host", f)
                    for f in os.listdir(os.path.join(ROOT, "02_cathode-08_modulerev_amd", "host"))]
    if (not force and os.path.exists(EXE)
            and os.path.getmtime(EXE) >= max(os.path.getmtime(d) for d in deps)):
        return EXE
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", SRC, "-o", EXE,
           "-L", CSRC, "-lcathode_hip", "-L", ORACLE, "-loracle", "-fopenmp",
           "-L/opt/rocm/lib", f"-Wl,-rpath,{CSRC}:{ORACLE}:/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return EXE
