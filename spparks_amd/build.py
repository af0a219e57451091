"""Build libspk_hip.so (HIP, gfx950) and the C++ front end in-tree.

    python -m spparks_amd.build

Outputs go to spparks_amd/lib/.  Always recompiles (it is the build check).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["spk_common.hip", "spk_plain.hip", "spk_batch.hip", "spk_serve.hip"]
FRONT_SOURCES = ["front/spk_front.cpp", "front/spk_dump_writer.cpp"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wextra", "-ffp-contract=off"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
    # bit-exact doubles: no a*b+c contraction anywhere (tagreset's rseed*IM+id)
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip():
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libspk_hip.so")
    tmp = out + ".tmp"
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
        run([hipcc(), *HIPCC_FLAGS, "-I" + INCLUDE, "-I" + CSRC, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    return out


def cxx():
    return os.environ.get("CXX") or shutil.which("g++") or "g++"


def build_front():
    """The spk_hip program and libspparks_hip.so (host C++ over the C ABI)."""
    inc = ["-I" + INCLUDE, "-I" + CSRC, "-I" + os.path.join(CSRC, "front")]
    link = ["-L" + LIBDIR, "-lspk_hip", "-Wl,-rpath,$ORIGIN"]
    objs = []
    for src in FRONT_SOURCES:
        obj = os.path.join(LIBDIR, os.path.basename(src).replace(".cpp", ".o"))
        run([cxx(), *CXX_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    lib = os.path.join(LIBDIR, "libspparks_hip.so")
    run([cxx(), *CXX_FLAGS, *inc, "-shared", "-o", lib + ".tmp", os.path.join(CSRC, "front", "spk_library.cpp"),
         *objs, *link])
    os.replace(lib + ".tmp", lib)
    exe = os.path.join(LIBDIR, "spk_hip")
    run([cxx(), *CXX_FLAGS, *inc, "-o", exe + ".tmp", os.path.join(CSRC, "front", "spk_main.cpp"), *objs, *link])
    os.replace(exe + ".tmp", exe)
    for o in objs:
        os.remove(o)
    return exe


def build_emulator():
    """Host-only emulator of the kernels' index math (tests/emul), built with
    AddressSanitizer for the CPU tests.  Test infrastructure."""
    emul = os.path.join(ROOT, "tests", "emul")
    src = os.path.join(emul, "spk_emul.hip")
    if not os.path.exists(src):
        return None
    out = os.path.join(emul, "spk_emul")
    run([hipcc(), "-x", "hip", "--offload-host-only", "-O1", "-g", "-std=c++20", "-ffp-contract=off",
         "-Xarch_host", "-fsanitize=address", "-I" + INCLUDE, "-I" + CSRC, "-o", out, src,
         "-L" + LIBDIR, "-lspk_hip", "-Wl,-rpath," + LIBDIR])
    return out


def build_abi_consumer():
    """tests/abi/consumer: a plain C99 caller of libspk_hip.so and
    libspparks_hip.so through their headers.  Test infrastructure."""
    src = os.path.join(ROOT, "tests", "abi", "consumer.c")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "abi", "consumer")
    run(["gcc", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-O1", "-I" + INCLUDE, "-o", out, src,
         "-L" + LIBDIR, "-lspparks_hip", "-lspk_hip", "-Wl,-rpath," + LIBDIR])
    return out


def build_refdecl():
    """tests/abi/refdecl: a C caller that declares the library entry points
    exactly as the reference's src/library.h does and links
    libspparks_hip.so.  Test infrastructure."""
    src = os.path.join(ROOT, "tests", "abi", "refdecl.c")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "abi", "refdecl")
    run(["gcc", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-O1", "-o", out, src,
         "-L" + LIBDIR, "-lspparks_hip", "-Wl,-rpath," + LIBDIR])
    return out


def build_dump_format_check():
    """tests/abi/dump_format: the dump writer's formatting against fprintf,
    host only.  Test infrastructure."""
    src = os.path.join(ROOT, "tests", "abi", "dump_format.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "abi", "dump_format")
    run([cxx(), *CXX_FLAGS, "-I" + os.path.join(CSRC, "front"), "-o", out, src,
         os.path.join(CSRC, "front", "spk_dump_writer.cpp")])
    return out


def build():
    build_hip()
    build_front()


if __name__ == "__main__":
    build()
    sys.exit(0)
