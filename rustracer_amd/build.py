"""Build recipe of the gfx950 library: hipcc -> rustracer_amd/librt_gfx950.so (in-tree).

Run `python -m rustracer_amd.build` (or __graft_entry__.build()).  Cross-compiles for
gfx950 on a machine without a GPU.
"""
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librt_gfx950.so")
SOURCES = ["api.hip", "lbvh.hip", "render.hip", "wavefront.hip", "adjoint.hip", "skin.hip"]
HEADERS = ["common.h", "lbvh.h", "traverse.h", "shade.h", "render.h", "path.h", "wavefront.h", "output.h", "skin.h",
           "brdf_dual.h", "adjoint.h"]
PUBLIC = [os.path.join(ROOT, "include", "rt", "rt.h"), os.path.join(ROOT, "include", "rt", "rt_layout.h")]

# Packed fp32 instructions are disabled: hipcc would otherwise form them from
# scalar float code, and some of their forms have given wrong results here.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-ffp-contract=on",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics",
         "-I", os.path.join(ROOT, "include")]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found; the gfx950 library cannot be built")


def build(verbose=False, jobs=4):
    """Compile every HIP source for gfx950 and link librt_gfx950.so.  Always rebuilds."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    t0 = time.time()
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [cc, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            time.sleep(0.05)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = LIB + ".tmp"
    subprocess.check_call([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s", flush=True)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
