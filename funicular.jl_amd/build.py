"""Builds libfunicular_hip.so for gfx950, in tree.

    python funicular.jl_amd/build.py

The library is plain hipcc output: every .hip/.cpp file under csrc/ is
compiled for gfx950 and linked into funicular_hip/libfunicular_hip.so next to
the Python host package that loads it. No PyTorch, no JIT cache.
"""
import os
import subprocess
import sys
import concurrent.futures

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "funicular_hip")
LIB = os.path.join(OUT_DIR, "libfunicular_hip.so")
OBJ_DIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # Packed FP32 instructions are kept out of the device code: the compiler
    # would otherwise form them from ordinary float arithmetic.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return sorted(
        os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp"))
    )


def compile_one(src, extra=(), obj_dir=None):
    obj = os.path.join(obj_dir or OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, *extra, "-c", src, "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless.
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in proc.stderr.splitlines() if noise not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{err}")
    return obj, err


def build(verbose=False, defines=(), out=None):
    """Compiles every source (with optional -D tuning defines) and links the
    library, by default in place of funicular_hip/libfunicular_hip.so."""
    obj_dir = OBJ_DIR if out is None else out + ".objs"
    os.makedirs(obj_dir, exist_ok=True)
    extra = [f"-D{d}" for d in defines]
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(lambda s: compile_one(s, extra, obj_dir), srcs))
    objs = [o for o, _ in results]
    if verbose:
        for (o, err) in results:
            if err.strip():
                print(err, file=sys.stderr)
    lib = LIB if out is None else out
    tmp = lib + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stderr}")
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
