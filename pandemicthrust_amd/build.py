"""Build recipe: the gfx950 library, the C++ host program, and (for tests)
the CPU oracle.

  lib/libpt_hip.so     hipcc --offload-arch=gfx950, IEEE single precision:
                       no contraction, correctly rounded divide, denormals kept
  bin/pandemicthrust   g++ host program linked against lib/libpt_hip.so
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HIP_SOURCES = ["pt_sim.hip", "pt_batch.hip", "pt_sort.hip", "pt_mlp.hip", "pt_table.hip", "pt_table_noalloc.hip"]


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build_hip():
    os.makedirs(os.path.join(HERE, "lib"), exist_ok=True)
    out = os.path.join(HERE, "lib", "libpt_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in HIP_SOURCES])
    os.replace(tmp, out)
    return out


def build_host():
    os.makedirs(os.path.join(HERE, "bin"), exist_ok=True)
    out = os.path.join(HERE, "bin", "pandemicthrust")
    src = [os.path.join(HERE, "host", f) for f in ("main.cpp", "pandemic_sim.cpp", "validation.cpp")]
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", out] + src +
         ["-L" + os.path.join(HERE, "lib"), "-lpt_hip", "-Wl,-rpath,$ORIGIN/../lib"])
    return out


def build_oracle():
    _run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    build_hip()
    build_host()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(HERE, "lib", "libpt_hip.so"))
