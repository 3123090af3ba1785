"""Builds parasail_amd/libparasail_hip.so in-tree.

C host layer (csrc/host/*.c) with gcc, HIP device layer (csrc/device/*.hip)
with hipcc for gfx950, linked into one shared library.  Objects go to
parasail_amd/_build/.

    python -m parasail_amd.build          # or parasail_amd.build.build()
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libparasail_hip.so")
DRIVER = os.path.join(HERE, "percall_omp")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PARASAIL_HIP_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    # packed fp32 VALU forms are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
C_FLAGS = ["-O2", "-fPIC", "-std=c11", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-D_POSIX_C_SOURCE=200112L", "-I" + os.path.join(HERE, "..", "include")]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "ignoring feature" not in l)
    if p.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    tasks = []
    for fn in sorted(os.listdir(os.path.join(CSRC, "host"))):
        if fn.endswith(".c"):
            src = os.path.join(CSRC, "host", fn)
            obj = os.path.join(BUILD, "host_" + fn[:-2] + ".o")
            tasks.append(["gcc"] + C_FLAGS + ["-c", src, "-o", obj])
    for fn in sorted(os.listdir(os.path.join(CSRC, "device"))):
        if fn.endswith(".hip"):
            src = os.path.join(CSRC, "device", fn)
            obj = os.path.join(BUILD, "dev_" + fn[:-4] + ".o")
            tasks.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
    jobs = jobs or min(8, os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        for out in ex.map(_run, tasks):
            if verbose and out.strip():
                print(out)
    objs = [t[-1] for t in tasks]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs +
         ["-lpthread", "-lz"])
    os.replace(tmp, LIB)
    # bench.py's per-call driver (OpenMP over targets, as parasail's own tests
    # call the per-pair functions); not part of the library
    _run(["gcc", "-O2", "-fopenmp", "-Wall", "-o", DRIVER,
          os.path.join(CSRC, "bench", "percall_omp.c"), "-L" + HERE, "-lparasail_hip",
          "-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
