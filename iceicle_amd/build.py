"""Build libiceicle_amd.so (HIP for gfx950) in-tree.

Every HIP translation unit is compiled with hipcc for gfx950 with packed
FP32 instructions disabled, then linked into one shared library next to this
file.  Used by __graft_entry__.build(); run directly with
`python -m iceicle_amd.build`.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libiceicle_amd.so")

SOURCES = ["mesh.cpp", "engine.cpp", "residual.hip", "stepper.hip", "jacobian.hip", "krylov.hip", "newton.cpp",
           "mlp.hip", "noalloc.hip", "mdg.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ICE_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++20", "-fPIC", "-Wall", "-Wno-unused-variable", "-Wno-unused-function",
          "-Wno-unused-but-set-variable"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src):
    out = os.path.join(OBJ, src + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [HIPCC, *COMMON, *DEVICE, "-c", path, "-o", out]
    else:
        cmd = [HIPCC, *COMMON, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-x", "c++", "-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and "remark" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{errs}")
    return out, errs


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        for out, errs in ex.map(_compile, SOURCES):
            objs.append(out)
            if verbose and errs.strip():
                print(errs, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed\n" + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
