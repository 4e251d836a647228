"""Builds the package's native libraries in-tree.

    libwhale_amd.so    C ABI + gfx950 kernels (hipcc --offload-arch=gfx950)
    libwhale_synth.so  seeded synthetic workload generator (host only, g++)

Run as `python whale.jl_amd/build.py` or through __graft_entry__.build().
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 VALU ops stay off in every HIP build of this project
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_main(out=None):
    out = out or os.path.join(HERE, "libwhale_amd.so")
    srcs = [os.path.join(CSRC, f) for f in ("kernels.hip", "capi.cpp", "host.cpp")]
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + srcs + ["-o", tmp, "-lpthread"])
    os.replace(tmp, out)
    return out


def build_synth(out=None):
    out = out or os.path.join(HERE, "libwhale_synth.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", os.path.join(CSRC, "synth.cpp"), "-o", tmp,
          "-lpthread"])
    os.replace(tmp, out)
    return out


def build_all():
    return build_main(), build_synth()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what or "main" in what:
        build_main()
    if "all" in what or "synth" in what:
        build_synth()
