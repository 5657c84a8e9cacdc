"""Builds libnebula_hip.so (the gfx950 kernels behind include/nebula_hip.h) in-tree with hipcc.

python -m nebulaviewer_amd.build          # or: python nebulaviewer_amd/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnebula_hip.so")
SOURCES = ["capi.hip", "hydro.hip", "rays.hip", "codec.hip", "pre.hip"]
ARCH = os.environ.get("NV_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off + IEEE division/sqrt: the kernels reproduce the CPU oracle bit for bit.
# -packed-fp32-ops off: packed FP32 instructions are not used on this target.
COMPILE = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def build(verbose=False):
    """Compiles each source to an object in parallel, then links the shared library (replaced atomically)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        procs.append((src, _run([hipcc] + COMPILE + ["-c", "-o", obj, os.path.join(CSRC, src)], verbose)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed.append(src)
    if failed:
        raise RuntimeError(f"hipcc failed on {failed} building {LIB}")
    tmp = LIB + ".tmp"
    link = _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs, verbose)
    out, _ = link.communicate()
    if link.returncode != 0:
        sys.stderr.write(out)
        raise RuntimeError(f"linking {LIB} failed ({link.returncode})")
    os.replace(tmp, LIB)
    return LIB


def build_shared(src, out, verbose=False):
    """Compiles one more HIP source into its own shared library with the same flags (the test helpers)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = out + ".tmp"
    p = _run([hipcc] + COMPILE + ["-shared", "-o", tmp, src], verbose)
    text, _ = p.communicate()
    if p.returncode != 0:
        sys.stderr.write(text)
        raise RuntimeError(f"hipcc failed on {src}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
