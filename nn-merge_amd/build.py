"""Build libnnmerge.so (HIP kernels + C ABI + C++ host classes) for gfx950.

    python nn-merge_amd/build.py          # compile every source, link the library

Objects go to nn-merge_amd/_build/, the library to nn-merge_amd/libnnmerge.so.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnnmerge.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", "-Wno-unused-value", "-I" + os.path.join(ROOT, "include")]
# Packed FP32 math stays disabled for every kernel.
HIPFLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]


def sources():
    hip = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp")) if os.path.isdir(HOST) else []
    return hip, cpp


def _compile(src):
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + COMMON + HIPFLAGS + ["-c", src, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + ["-x", "c++", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    hip, cpp = sources()
    objs = []
    errors = []
    jobs = min(8, max(1, len(hip) + len(cpp)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        for src, obj, rc, log in ex.map(_compile, hip + cpp):
            if rc != 0:
                errors.append((src, log))
            else:
                objs.append(obj)
                if verbose and log.strip():
                    print(log)
    if errors:
        for src, log in errors:
            sys.stderr.write(f"=== {src}\n{log}\n")
        raise RuntimeError("libnnmerge build failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("libnnmerge link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
