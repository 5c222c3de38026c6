"""Builds the native parts of care_amd in-tree (care_amd/_build/).

  libcare_amd.so   HIP engine for gfx950 (care_amd/csrc/*.hip, C-ABI in include/care_amd.h)
  libcare_host.so  C++17 host: care-gpu command line and read packing (include/care_host.h)
  care-gpu         the command-line corrector
  readsim          seeded read simulator / base-level evaluator
"""

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # IEEE float arithmetic exactly as written: no contraction into FMA,
    # correctly rounded division and square root
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

HIP_SOURCES = ["api.hip", "minhash.hip", "correct.hip", "mlp.hip", "store.hip"]


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "is not a recognized feature" not in l
                    and "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_engine():
    os.makedirs(BUILD, exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", s) for s in HIP_SOURCES]
    out = os.path.join(BUILD, "libcare_amd.so")
    # intra-library calls bind to this library even if another one exporting
    # the same C-ABI is loaded into the process first
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-Wl,-Bsymbolic", "-o", out + ".tmp"] + srcs)
    os.replace(out + ".tmp", out)
    return out


def build_host():
    os.makedirs(BUILD, exist_ok=True)
    lib = os.path.join(BUILD, "libcare_host.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", lib + ".tmp",
          os.path.join(HERE, "host", "care_host.cpp"), os.path.join(HERE, "host", "storage_files.cpp"),
          "-L" + BUILD, "-lcare_amd",
          "-Wl,-rpath,$ORIGIN", "-lz", "-lpthread"])
    os.replace(lib + ".tmp", lib)
    exe = os.path.join(BUILD, "care-gpu")
    _run(["g++", "-O2", "-std=c++17", "-o", exe + ".tmp", os.path.join(HERE, "host", "care_gpu_main.cpp"),
          "-L" + BUILD, "-lcare_host", "-lcare_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(exe + ".tmp", exe)
    return exe


def build_sim():
    os.makedirs(BUILD, exist_ok=True)
    exe = os.path.join(BUILD, "readsim")
    _run(["g++", "-O2", "-std=c++17", "-o", exe + ".tmp", os.path.join(HERE, "sim", "readsim.cpp")])
    os.replace(exe + ".tmp", exe)
    return exe


def build_all():
    build_engine()
    build_host()
    build_sim()


if __name__ == "__main__":
    build_all()
    print("built into", BUILD)
