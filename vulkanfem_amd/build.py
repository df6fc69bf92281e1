"""Builds the HIP library vulkanfem_amd/libvfem.so for gfx950 with hipcc, and the
command-line solver vulkanfem_amd/vfem_cli (C++, over include/vfem.hpp) against it.

    python -m vulkanfem_amd.build

Both are built in-tree (they travel with the source tree); nothing is
installed.  Every HIP source is compiled with packed FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvfem.so")
CLI = os.path.join(HERE, "vfem_cli")

SOURCES = ["vfem_setup.cpp", "vfem_setup_device.hip", "vfem_device.hip", "vfem_unet.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=2):
    """Compile every source (always -- this is the build check) and link libvfem.so."""
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    with tempfile.TemporaryDirectory() as tmp:
        objs = []
        procs = []
        for src in SOURCES:
            obj = os.path.join(tmp, os.path.splitext(src)[0] + ".o")
            lang = ["-x", "hip"] if src.endswith(".hip") else []
            cmd = [hipcc()] + HIP_FLAGS + inc + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
            objs.append(obj)
        failed = False
        for src, p in procs:
            out, _ = p.communicate()
            if p.returncode != 0 or (verbose and out.strip()):
                sys.stderr.write(f"--- {src}\n{out}\n")
            failed |= p.returncode != 0
        if failed:
            raise RuntimeError("hipcc failed")
        tmp_lib = LIB + ".tmp"
        cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp_lib] + objs + ["-lpthread"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp_lib, LIB)
    build_cli(verbose)
    return LIB


def build_cli(verbose=False):
    """The command-line drop-in for the reference's solver program: host C++ only,
    linked against libvfem.so, found next to it at run time."""
    cxx = os.environ.get("CXX") or shutil.which("g++") or "g++"
    tmp = CLI + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "vfem_cli.cpp"),
           "-L", HERE, "-lvfem", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB)
