"""Build the apyib_amd HIP library for gfx950.

Compiles every ``csrc/*.hip`` translation unit with hipcc into the in-tree
shared library ``apyib_amd/libapyib_amd.so`` (loaded with ctypes by
``apyib_amd._lib``).  Always rebuilds; objects are compiled in parallel.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libapyib_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    # Packed FP32 instructions stay disabled for every kernel in this library.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in ordinary VGPRs: no copies between the VGPR and
    # accumulation register files around the GEMM main loops.
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wno-unused-result",
]


def _compile(src, obj):
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{errs}")
    return errs


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = []
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        for f in srcs:
            obj = os.path.join(objdir, f[:-4] + ".o")
            jobs.append((f, ex.submit(_compile, os.path.join(CSRC, f), obj), obj))
        objs = []
        for f, fut, obj in jobs:
            msg = fut.result()
            if verbose and msg.strip():
                print(msg, file=sys.stderr)
            objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
