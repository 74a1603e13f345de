"""Build libfdtd3d.so in-tree: C host layer (gcc) + HIP kernels for gfx950 (hipcc).

    python 3d_fdtd_amd/build.py            # -> 3d_fdtd_amd/lib/libfdtd3d.so

Always recompiles (it is the "does it build" check); objects go to 3d_fdtd_amd/_build/.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(PKG, "_build")
LIB_DIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIB_DIR, "libfdtd3d.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", "-Wno-unused-result", "-I" + INC]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu99", "-Wall", "-Wno-unused-result", "-I" + INC]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    err = "\n".join(ln for ln in (r.stderr or "").splitlines() if noise not in ln)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + err)
    return err


def sources():
    hip = sorted(glob.glob(os.path.join(PKG, "csrc", "hip", "*.hip")))
    c = sorted(glob.glob(os.path.join(PKG, "csrc", "host", "*.c")))
    return hip, c


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    hip, c = sources()
    jobs = []
    for s in hip:
        o = os.path.join(OBJ, os.path.basename(s) + ".o")
        jobs.append([HIPCC, "-c", s, "-o", o] + HIP_FLAGS)
    for s in c:
        o = os.path.join(OBJ, os.path.basename(s) + ".o")
        jobs.append(["gcc", "-c", s, "-o", o] + C_FLAGS)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for msg in ex.map(_run, jobs):
            if verbose and msg.strip():
                print(msg)
    objs = [j[4] for j in jobs]
    tmp = os.path.join(LIB_DIR, "libfdtd3d.partial.so")
    _run([HIPCC, "-shared", "-o", tmp, f"--offload-arch={ARCH}"] + objs + ["-lm"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
