"""Builds libmipt_hip.so for gfx950 in-tree (hipcc, no JIT cache)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmipt_hip.so")
SOURCES = ["plain.hip", "batch.hip", "entropy.hip", "eigen.hip", "decoder.hip", "driver.cpp", "probe_info.cpp"]
BIN = os.path.join(HERE, "bin")
TOOLS = {"sim_tmi": os.path.join(HERE, "host", "sim_tmi_main.cpp"), "mipt": os.path.join(HERE, "host", "mipt_main.cpp"),
         "entropy": os.path.join(HERE, "host", "entropy_main.cpp"),
         "mipt_probed": os.path.join(HERE, "host", "probed_main.cpp"),
         "free_energy": os.path.join(HERE, "host", "free_energy_main.cpp")}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    # packed FP32 math stays off (v_pk_* FP32 forms)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-I/opt/rocm/include",
]
LIBS = ["-shared", "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def build(verbose: bool = False) -> str:
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = False
    for cmd, p in procs:
        out = p.communicate()[0].decode(errors="replace")
        lines = [l for l in out.splitlines() if "is not a recognized feature" not in l]
        if p.returncode != 0:
            failed = True
            sys.stderr.write(" ".join(cmd) + "\n" + "\n".join(lines) + "\n")
        elif verbose and lines:
            sys.stderr.write("\n".join(lines) + "\n")
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = OUT + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", *objs, *LIBS, "-o", tmp], check=True)
    os.replace(tmp, OUT)
    # host executables over the C ABI (plain C++, found next to the library)
    os.makedirs(BIN, exist_ok=True)
    for name, src in TOOLS.items():
        subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), src, "-L" + HERE,
                        "-lmipt_hip", "-Wl,-rpath,$ORIGIN/..", "-o", os.path.join(BIN, name)], check=True)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
