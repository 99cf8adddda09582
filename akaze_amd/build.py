"""Build recipe of libakaze_amd.so (HIP, gfx950) and of the CPU oracle.

Run as `python -m akaze_amd.build` or through __graft_entry__.build().  Every
translation unit is compiled in parallel with hipcc, then linked in-tree into
akaze_amd/libakaze_amd.so.

Flags:
  -ffp-contract=off   no FMA contraction: float expressions evaluate exactly as
                      written, so kernels reproduce the CPU oracle bit for bit;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libakaze_amd.so")
ARCH = os.environ.get("AKZ_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "akz_scale_space.hip",
    "akz_detect.hip",
    "akz_describe.hip",
    "akz_match.hip",
    "akz_proposal.hip",
    "akz_synth_dev.hip",
    "akz_capi.hip",
    "akz_synth.cpp",
    "akz_utils.cpp",
    "akz_mldb_table.cpp",
    "AKAZE.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I", os.path.join(REPO, "include"), "-I", CSRC]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-fgpu-rdc" if False else "-fno-gpu-rdc"]


def _compile(src):
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + DEVICE + COMMON + ["-c", path, "-o", out]
    else:
        cmd = [HIPCC] + COMMON + ["-D__HIP_PLATFORM_AMD__", "-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


APPS = ["akaze_features", "akaze_match"]
BIN = os.path.join(HERE, "bin")


def build_apps(verbose=True):
    """The GUI-free command-line programs, linked against the library in-tree
    (rpath $ORIGIN/..)."""
    os.makedirs(BIN, exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    for app in APPS:
        out = os.path.join(BIN, app)
        cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(REPO, "include"),
               os.path.join(HERE, "apps", app + ".cpp"), "-o", out + ".tmp", "-L", HERE, "-l:libakaze_amd.so",
               "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("app build failed: %s\n%s" % (" ".join(cmd), r.stderr))
        os.replace(out + ".tmp", out)
        if verbose:
            print("built", out)


def build_oracle(verbose=True):
    r = subprocess.run(["make", "-C", os.path.join(REPO, "oracle")], capture_output=True, text=True)
    if verbose:
        print(r.stdout[-2000:], r.stderr[-2000:])
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_apps(verbose)
    return LIB


if __name__ == "__main__":
    build()
    if "--no-oracle" not in sys.argv:
        build_oracle()
