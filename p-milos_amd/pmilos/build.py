"""Builds libpmilos_hip.so (the HIP kernels + C ABI) in-tree for gfx950.

    python p-milos_amd/pmilos/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
LIB = os.path.join(HERE, "libpmilos_hip.so")
SOURCES = ["pm_kernels.hip", "pm_server.hip"]

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 instructions (v_pk_*_f32) are generated
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


CLI_DIR = os.path.join(os.path.dirname(HERE), "cli")
IO_LIB = os.path.join(HERE, "libpmilos_io.so")
CLI_BIN = os.path.join(HERE, "pmilos_gpu")
CXX_FLAGS = ["-O2", "-std=c++17", "-Wall", "-Wno-unused-result"]


def build_cli(verbose=False):
    """The command line tool (host C++ on the C ABI) and its file-format
    library, which the CPU tests load on their own."""
    cxx = os.environ.get("CXX", "g++")
    io_src = os.path.join(CLI_DIR, "pm_io.cpp")
    cmds = [
        [cxx] + CXX_FLAGS + ["-fPIC", "-shared", "-o", IO_LIB + ".tmp", io_src],
        [cxx] + CXX_FLAGS + ["-o", CLI_BIN + ".tmp", os.path.join(CLI_DIR, "pmilos_cli.cpp"), io_src,
                             "-L" + HERE, "-lpmilos_hip", "-Wl,-rpath,$ORIGIN", "-lpthread"],
    ]
    for cmd, out in zip(cmds, (IO_LIB, CLI_BIN)):
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stderr)
            raise RuntimeError("%s failed (%d)" % (cxx, res.returncode))
        if verbose and res.stderr.strip():
            print(res.stderr)
        os.replace(out + ".tmp", out)
    return CLI_BIN


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    errs = [l for l in res.stderr.splitlines()
            if l.strip() and "packed-fp32-ops" not in l and "warning: unknown" not in l]
    if res.returncode != 0:
        sys.stderr.write(res.stderr)
        raise RuntimeError("hipcc failed (%d)" % res.returncode)
    if verbose and errs:
        print("\n".join(errs))
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
