"""Build recipe for librnpx.so (the gfx950 HIP library behind include/rnpx.h).

Every source is compiled with hipcc for gfx950 only, in parallel, into
rnp2_amd/_build/, and linked in-tree as rnp2_amd/librnpx.so so that the
library travels with the source tree.

    python -m rnp2_amd.build        # or rnp2_amd.build.build()
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librnpx.so")
ARCH = "gfx950"

SOURCES = [
    "rnpx_panel.hip",
    "rnpx_cqr.hip", "rnpx_trsm.hip", "rnpx_lu.hip",
    "rnpx_tfactor.hip",
    "rnpx_gemm.hip",
    "rnpx_dgemm.hip",
    "rnpx_aux.hip",
    "rnpx_batched.hip",
    "rnpx_mixed.hip",
    "rnpx_noalloc.hip",
    "rnpx_exact.hip",
    "rnpx_zgemm.hip",
    "rnpx_zcqr.hip",
    "rnpx_gemv.hip",
    "rnpx_qr.cpp", "rnpx_qlrq.cpp", "rnpx_hess.cpp", "rnpx_lu_driver.cpp",
    "rnpx_lq.cpp",
    "rnpx_ls.cpp",
    "rnpx_serve.cpp",
    "rnpx_api.cpp",
    "rnpx_tune.cpp",
]

HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-lambda-capture",
    # Packed FP32 VALU forms are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def _compile(src, verbose):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), "-x", "hip", *HIPFLAGS, "-I", os.path.join(HERE, "..", "include"), "-c", path, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip() and verbose:
        print(r.stderr, file=sys.stderr)
    return obj


def build(verbose=False, jobs=None):
    """Compile every source for gfx950 and link librnpx.so in place."""
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-ldl"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
