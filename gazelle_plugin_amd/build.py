"""Builds libgazelle_amd.so (HIP, gfx950) in-tree with hipcc, and over it
libgazelle_amd_arrow.so, the C++ operator interfaces on Arrow C++ (g++; the
Arrow C++ headers and library come with pyarrow)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgazelle_amd.so")

SOURCES = ["gz_context.cpp", "gz_split.hip", "gz_ipc_read.hip", "gz_agg.hip", "gz_keys.hip", "gz_join.hip", "gz_serve.hip",
           "gz_stage.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled: hipcc otherwise forms v_pk_* from
# ordinary float code.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-value"]


ARROW_LIB = os.path.join(HERE, "libgazelle_amd_arrow.so")
ARROW_SOURCES = [os.path.join(HERE, "cpp", "gz_arrow.cc")]


def pyarrow_dir() -> str:
    import pyarrow

    return os.path.dirname(pyarrow.__file__)


def _run(cmd, verbose, what):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(what + " failed:\n" + err)
    if verbose and err.strip():
        print(err, file=sys.stderr)


def build_arrow_layer(verbose: bool = False) -> str:
    """libgazelle_amd_arrow.so: host C++ over libgazelle_amd.so and libarrow
    (Arrow C++ 25 needs C++20)."""
    pa = pyarrow_dir()
    cmd = (["g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-Wall", "-D__HIP_PLATFORM_AMD__",
            "-I", os.path.join(HERE, "..", "include"), "-I", os.path.join(pa, "include"), "-I", "/opt/rocm/include"]
           + ARROW_SOURCES
           + ["-o", ARROW_LIB + ".tmp", "-L", HERE, "-lgazelle_amd", "-L", pa, "-l:libarrow.so.2500",
              "-L", "/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + pa,
              "-Wl,-rpath,/opt/rocm/lib"])
    _run(cmd, verbose, "g++ (Arrow layer)")
    os.replace(ARROW_LIB + ".tmp", ARROW_LIB)
    return ARROW_LIB


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(HERE, "..", "include")] + srcs + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + err)
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(LIB + ".tmp", LIB)
    build_arrow_layer(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
