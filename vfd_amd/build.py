"""Builds the HIP extension (libvfd_hip.so, gfx950) in-tree with hipcc.

    python -m vfd_amd.build            # or: python vfd_amd/build.py

Every source is compiled with -ffp-contract=off (fused multiply-adds only where the code writes fmaf) and with
packed FP32 instructions disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvfd_hip.so")
SOURCES = ["search.hip", "boundary.hip", "dfsph.hip", "adjoint.hip", "sim.cpp", "scene_io.cpp", "mesh.cpp"]
ARCH = os.environ.get("VFD_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags():
    return [
        "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
        "--offload-arch=" + ARCH,
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
        "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    ]


def _run(cmd):
    """Runs a compiler command; drops the host pass's note that it does not know the packed-fp32 feature."""
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if p.returncode != 0:
        raise subprocess.CalledProcessError(p.returncode, cmd)


def build(verbose=False, out=None, defines=()):
    """Builds LIB (or `out`); `defines` are compile-time launch-shape choices such as "VFD_LIST_CHUNK=16" (see
    vfd_amd.autotune)."""
    cc = hipcc()
    if out is None:
        from vfd_amd.autotune import chosen_defines
        defines = list(chosen_defines()) + list(defines)
    out = out or LIB
    objdir = os.path.join(os.path.dirname(out), "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [cc] + flags() + ["-D" + d for d in defines] + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        objs.append(obj)
    tmp = out + ".tmp"
    cmd = [cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
