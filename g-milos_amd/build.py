"""Build recipe for the gfx950 inversion library (``lib/libgmilos.so``).

Compiles the HIP kernels and the C ABI with hipcc for gfx950.  Packed FP32
instructions are disabled for all device code.  Used by the package's
``build()`` and by ``__graft_entry__.build``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib", "libgmilos.so")
SOURCES = ["csrc/gm_kernels.hip", "csrc/gm_capi.cpp", "csrc/gm_setup.cpp", "csrc/gm_mlp.hip",
           "csrc/gm_mlp_host.cpp", "csrc/gm_server.cpp", "csrc/gm_guided.hip",
           "csrc/gm_noalloc.hip"]
IO_SOURCES = ["host/gm_io.cpp", "host/gm_fits.cpp"]
IO_LIB = os.path.join(HERE, "lib", "libgmilos_io.so")
CLI_SOURCES = ["host/gmilos_cli.cpp"]
CLI = os.path.join(HERE, "bin", "gmilos")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    proc = subprocess.run(cmd, cwd=HERE, stdout=subprocess.PIPE,
                          stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "nodiscard" not in l)
    if proc.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


def build_lib(verbose=False):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Wno-unused-result", "-Wno-unused-value",
           "-I" + os.path.join(ROOT, "include"), "-Icsrc"] + DEVICE_FLAGS + SOURCES + ["-o", tmp]
    out = _run(cmd)
    os.replace(tmp, LIB)
    if verbose and out.strip():
        print(out)
    return LIB


def build_io(verbose=False):
    """Host-only file-format library (.mtrol, .mod, .grid, .psf, .per, FITS)."""
    os.makedirs(os.path.dirname(IO_LIB), exist_ok=True)
    tmp = IO_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
          "-I" + os.path.join(ROOT, "include")] + IO_SOURCES + ["-o", tmp])
    os.replace(tmp, IO_LIB)
    return IO_LIB


def build_cli(verbose=False):
    """The ``gmilos <file>.mtrol`` command (host C++ over both libraries)."""
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
          "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROCM, "include")] +
         CLI_SOURCES + ["-L" + os.path.dirname(LIB), "-lgmilos", "-lgmilos_io",
                        "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
                        "-Wl,-rpath,$ORIGIN/../lib:" + os.path.join(ROCM, "lib"), "-o", tmp])
    os.replace(tmp, CLI)
    return CLI


def build(verbose=False):
    build_lib(verbose)
    build_io(verbose)
    build_cli(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
