"""Build libhorsepower_amd.so (all HIP sources, gfx950) in-tree with hipcc.

Usage: python -m horsepower_amd.build   (or horsepower_amd.build.build())
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhorsepower_amd.so")
ARCH = os.environ.get("HPA_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: FP64 arithmetic is rounded per operation as in the reference's C
# (no fused multiply-add), so products such as price*(1-discount) match bit for bit.
# The packed-fp32 target feature is disabled for every HIP translation unit.
HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-variable", "-Wno-unused-function",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.exists(cand) or cand == "hipcc"):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [_hipcc(), *HIPFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread", "-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_c_driver(verbose)
    return LIB


ROOT = os.path.dirname(HERE)
DRIVER_SRC = os.path.join(ROOT, "tests", "c_driver", "driver.c")
DRIVER = os.path.join(ROOT, "tests", "c_driver", "hpa_c_driver")


def build_c_driver(verbose=False):
    """The C caller the GPU tests run (tests/c_driver/driver.c): plain gcc against the public
    header, linked to the library with a run path relative to the binary."""
    if not os.path.exists(DRIVER_SRC):
        return None
    cmd = ["gcc", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), DRIVER_SRC, "-L", HERE, "-lhorsepower_amd",
           "-Wl,-rpath,$ORIGIN/../../horsepower_amd", "-o", DRIVER]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C driver build failed:\n{r.stderr[-4000:]}")
    if verbose:
        print(f"built {DRIVER}")
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
