"""Builds the engine in-tree: libgsofa.so (HIP for gfx950 + host C++) and the gsofa_cli binary.

Run as ``python -m gsofatpds_amd.build`` or through ``__graft_entry__.build()``; every call
recompiles. Nothing here is needed at run time.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgsofa.so")
CLI = os.path.join(HERE, "gsofa_cli")
CONVERT = os.path.join(HERE, "gsofa_convert")

SOURCES = ["kernels.hip", "engine.cpp", "api.cpp", "server.cpp", "host.cpp", "gen.cpp", "supernodal.cpp", "noalloc.hip",
           "convert.cpp"]
# -packed-fp32-ops: no v_pk_*_f32 generation on this pool (the engine has no fp32 math anyway).
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=CSRC)


def build(verbose=True):
    from concurrent.futures import ThreadPoolExecutor
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(CSRC, "_build", os.path.splitext(src)[0] + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        cmds.append([hipcc, *HIP_FLAGS, *lang, "-c", src, "-o", obj])
        objs.append(obj)
    jobs = max(1, min(len(cmds), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(jobs) as pool:  # translation units compile independently
        for f in [pool.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc, *HIP_FLAGS, "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    for src, exe in (("gsofa_cli.cpp", CLI), ("gsofa_convert.cpp", CONVERT)):
        _run([hipcc, "-O2", "-std=c++17", src, "-o", exe, "-L" + HERE, "-lgsofa", "-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
