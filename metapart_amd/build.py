"""Build libmetapart_amd.so (HIP, gfx950), the command-line tools and the
synthetic-data library, in-tree.

    python -m metapart_amd.build            # everything
    python -m metapart_amd.build --lib      # library only

Outputs (git-ignored, shipped with the tree):
    metapart_amd/libmetapart_amd.so
    metapart_amd/libmpsynth.so
    metapart_amd/bin/{genFastqIdx,genHistograms_metapart,genHistograms_metapartmin,
                      metapart,metapartmin,viewIdx,mp_synth}
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
TOOLS = os.path.join(PKG, "tools")
BUILD = os.path.join(PKG, "_build")
BIN = os.path.join(PKG, "bin")
LIB = os.path.join(PKG, "libmetapart_amd.so")
SYNTH = os.path.join(PKG, "libmpsynth.so")

ARCH = os.environ.get("MP_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed fp32 instructions stay off: no kernel here needs them
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HIP_SRCS = ["mp_util.hip", "mp_parse.hip", "mp_kmers.hip", "mp_sort.hip", "mp_graph.hip", "mp_writer.hip",
            "mp_api.hip", "mp_noalloc.hip"]
HOST_SRCS = ["mp_host.cpp", "mp_server.cpp"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd), r.stdout[-6000:]))
    return r.stdout


def build_lib(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs_list = []
    for s in HIP_SRCS:
        o = os.path.join(BUILD, s.replace(".hip", ".o"))
        jobs_list.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    for s in HOST_SRCS:
        o = os.path.join(BUILD, s.replace(".cpp", ".o"))
        jobs_list.append(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", "-I/opt/rocm/include",
                          "-D__HIP_PLATFORM_AMD__", "-c", os.path.join(CSRC, s), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, jobs_list))
    objs = [os.path.join(BUILD, s.replace(".hip", ".o")) for s in HIP_SRCS] + \
           [os.path.join(BUILD, s.replace(".cpp", ".o")) for s in HOST_SRCS]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-pthread", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_synth() -> str:
    tmp = SYNTH + ".tmp"
    _run(["gcc", "-O3", "-fopenmp", "-fPIC", "-shared", "-o", tmp, os.path.join(CSRC, "mp_synth.c"), "-lm"])
    os.replace(tmp, SYNTH)
    os.makedirs(BIN, exist_ok=True)
    _run(["gcc", "-O3", "-fopenmp", "-o", os.path.join(BIN, "mp_synth"), os.path.join(TOOLS, "mp_synth_main.c"),
          os.path.join(CSRC, "mp_synth.c"), "-lm"])
    return SYNTH


def build_tools() -> None:
    os.makedirs(BIN, exist_ok=True)
    common = ["gcc", "-O2", "-std=gnu11", "-Wall", "-Wno-unused-function"]
    link = ["-L" + PKG, "-lmetapart_amd", "-Wl,-rpath,$ORIGIN/.."]
    specs = [("genFastqIdx", "genFastqIdx.c", []), ("viewIdx", "viewIdx.c", []),
             ("genHistograms_metapart", "genHistograms.c", []),
             ("genHistograms_metapartmin", "genHistograms.c", ["-DMIN3"]),
             ("metapart", "metapart.c", []), ("metapartmin", "metapart.c", ["-DMIN3"])]
    for name, src, defs in specs:
        _run(common + defs + ["-o", os.path.join(BIN, name), os.path.join(TOOLS, src)] + link)


def build_oracle(with_ref: bool = True) -> None:
    """The CPU checker: oracle/liboracle.so always; oracle/_ref when the reference sources exist here."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "oracle"])
    if with_ref and os.path.isdir("/root/reference/src"):
        _run(["make", "-C", odir, "all"])


def build_all(jobs: int = 8) -> None:
    build_lib(jobs)
    build_synth()
    build_tools()
    build_oracle()


if __name__ == "__main__":
    if "--lib" in sys.argv:
        print(build_lib())
    else:
        build_all()
        print("built", LIB)
