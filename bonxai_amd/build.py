"""Build recipe of the native libraries (run by __graft_entry__.build()).

Outputs, all in-tree so they travel with the source tree:
  bonxai_amd/libbonxai_gpu.so   HIP library behind include/bonxai_gpu.h (gfx950)
  bonxai_amd/libbonxai_probes.so  measurement probes behind include/bonxai_gpu_probes.h
  bonxai_amd/libbxscenes.so     seeded synthetic workloads (C)
  bonxai_amd/bin/*              C++ drop-in programs built on include/bonxai*/
Each step compiles every time it is called.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ARCH = os.environ.get("BXG_OFFLOAD_ARCH", "gfx950")

# Every HIP translation unit: gfx950 only, FP64 contraction off (the scan front
# end must round like the reference's x86-64 double code), packed-FP32
# instructions off.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value",
    "-Wno-unused-result",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    proc = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return proc.stdout


def build_gpu_library():
    out = os.path.join(HERE, "libbonxai_gpu.so")
    tmp = out + ".tmp"
    _run([_hipcc(), *HIP_FLAGS, "-shared", "-I", os.path.join(ROOT, "include"),
          os.path.join(HERE, "csrc", "bxg_api.hip"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_probe_library():
    """The measurement probes, in a library of their own beside the map
    library (they run on a map context's stream)."""
    out = os.path.join(HERE, "libbonxai_probes.so")
    tmp = out + ".tmp"
    _run([_hipcc(), *HIP_FLAGS, "-shared", "-I", os.path.join(ROOT, "include"),
          os.path.join(HERE, "csrc", "bxg_probes.hip"), "-L", HERE, "-lbonxai_gpu",
          "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_scenes():
    out = os.path.join(HERE, "libbxscenes.so")
    _run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wall",
          os.path.join(HERE, "scenes", "scenes.c"), "-o", out, "-lm"])
    return out


def build_dropin_programs():
    """C++ programs that use the drop-in headers (Bonxai::VoxelGrid<T>,
    Bonxai::ProbabilisticMap) exactly as reference callers do."""
    src_dir = os.path.join(HERE, "cpp")
    bin_dir = os.path.join(HERE, "bin")
    os.makedirs(bin_dir, exist_ok=True)
    outs = []
    if not os.path.isdir(src_dir):
        return outs
    for name in sorted(os.listdir(src_dir)):
        if not name.endswith(".cpp"):
            continue
        out = os.path.join(bin_dir, name[:-4])
        _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-Wextra",
              "-I", os.path.join(ROOT, "include"), os.path.join(src_dir, name),
              "-L", HERE, "-lbonxai_gpu", "-Wl,-rpath,$ORIGIN/..", "-o", out])
        outs.append(out)
    return outs


def build_all():
    outs = [build_gpu_library(), build_probe_library(), build_scenes()]
    outs += build_dropin_programs()
    return outs


if __name__ == "__main__":
    for p in build_all():
        print(p)
