"""Build libxss_amd.so (hand-written HIP kernels for gfx950 + the C/C++ ABI).

    python x86-simd-sort_amd/build.py              # full rebuild
    python x86-simd-sort_amd/build.py --incremental
    python x86-simd-sort_amd/build.py --out DIR -D XSS_PASS_ITEMS_C1=12   # a tuning variant

Tuning knobs (-D, compile-time): XSS_PASS_BLOCK_Cn, XSS_PASS_ITEMS_Cn and
XSS_SCATTER_WAVES_Cn (n = 0..3, the item-width classes of xss_kernels.hpp)
set the tile shape and register budget of the scatter passes.  --out builds into DIR/_build and DIR/libxss_amd.so
instead of the package directory.

Every translation unit under csrc/ is compiled with hipcc for gfx950 into
_build/, then linked into x86-simd-sort_amd/libxss_amd.so (in-tree, so the
library travels with the source tree).
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libxss_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off",
    "-Wno-unused-result",
]
EXTRA = []  # -D tuning knobs


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, f))
    return out


def headers():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))] + [
        os.path.join(HERE, "..", "include", f) for f in os.listdir(os.path.join(HERE, "..", "include"))
    ]


def obj_of(src):
    return os.path.join(BUILD, os.path.basename(src) + ".o")


def compile_one(src):
    obj = obj_of(src)
    cmd = [HIPCC] + COMMON + EXTRA + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(incremental=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    hdr_mtime = max(os.path.getmtime(h) for h in headers())
    todo = []
    for s in srcs:
        o = obj_of(s)
        if (not incremental or not os.path.exists(o)
                or os.path.getmtime(o) < max(os.path.getmtime(s), hdr_mtime)):
            todo.append(s)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, todo))
    objs = [obj_of(s) for s in srcs]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_test_libs():
    """Test libraries built the way a caller builds against include/: the
    header-only device object_qsort instantiated in tests/cpp/object_gpu.hip
    (linked to libxss_amd.so, found through its rpath)."""
    repo = os.path.dirname(HERE)
    src = os.path.join(repo, "tests", "cpp", "object_gpu.hip")
    out = os.path.join(repo, "tests", "cpp", "libxss_objtest.so")
    cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=off",
           "-I", os.path.join(repo, "include"), src, "-o", out, "-L", HERE, "-lxss_amd",
           "-Wl,-rpath,$ORIGIN/../../x86-simd-sort_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"test library build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-D", dest="defines", action="append", default=[])
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    EXTRA[:] = [f"-D{d}" for d in a.defines]
    if a.out:
        BUILD = os.path.join(os.path.abspath(a.out), "_build")
        LIB = os.path.join(os.path.abspath(a.out), "libxss_amd.so")
    print(build(a.incremental, a.j))
