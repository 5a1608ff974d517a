"""Build libaether_hip.so (gfx950) in-tree.

Every .hip under aether_amd/csrc is compiled by hipcc for gfx950 with
contraction off (bit parity with the CPU restatement) and packed FP32
disabled, then linked into aether_amd/lib/libaether_hip.so.
Usage: python -m aether_amd.build  (or aether_amd.build.build()).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libaether_hip.so")
API_CHECK = os.path.join(LIB_DIR, "api_check")
RESIDENCY_CHECK = os.path.join(LIB_DIR, "residency_check")
FLAG_CHECK = os.path.join(LIB_DIR, "flag_check")
EXPR_DEVICE_CHECK = os.path.join(LIB_DIR, "expr_device_check")
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AETH_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++20", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


JIT_HEADERS = ["dmath.hpp", "rng.hpp", "accum.hpp", "units.hpp"]


def write_jit_prelude() -> str:
    """Embed the device headers (and the C ABI's tape structs) as one source string for hipRTC.

    The runtime compiler lowers each distinct expression tape to its own kernel; it includes
    exactly the device code the ahead-of-time kernels use, so both produce the same bits."""
    import re

    head = ("typedef unsigned char uint8_t; typedef unsigned short uint16_t; typedef unsigned int uint32_t;\n"
            "typedef unsigned long long uint64_t; typedef int int32_t; typedef long long int64_t;\n"
            "typedef decltype(sizeof(0)) size_t;\n")
    parts = [head]
    hdr = open(os.path.join(HERE, "..", "include", "aether_hip.h")).read()
    parts.append(re.sub(r"^#include .*$", "", hdr, flags=re.M))
    for f in JIT_HEADERS:
        text = open(os.path.join(CSRC, f)).read()
        text = re.sub(r"^#pragma once\s*$", "", text, flags=re.M)
        text = re.sub(r"^#include .*$", "", text, flags=re.M)
        parts.append(text)
    body = "\n".join(parts)
    assert ")AETHJIT" not in body
    out = os.path.join(BUILD, "jit_prelude.inc")
    with open(out, "w") as fh:
        fh.write('static const char kJitPrelude[] = R"AETHJIT(' + body + ')AETHJIT";\n')
    return out


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-I", BUILD, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *FLAGS, "-I", BUILD, "-x", "hip", "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    write_jit_prelude()
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread", "-lhiprtc"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    build_api_check()
    core = build_dsc_core()
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources, {API_CHECK} and {core}")
    return LIB


def _host_driver(name: str, out: str) -> str:
    src = os.path.join(ROOT, "tests", "cpp", name)
    cmd = ["g++", "-std=c++20", "-O2", "-ffp-contract=off", "-Wall", "-Wextra", "-Werror",
           "-I", os.path.join(ROOT, "include"), src, "-o", out + ".tmp", "-L", LIB_DIR, "-laether_hip",
           "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{res.stderr}")
    os.replace(out + ".tmp", out)
    return out


def build_api_check() -> str:
    """The C++ front end's test drivers (tests/cpp/api_check.cpp, tests/cpp/residency_check.cpp): host-only
    g++ builds against include/aether/*.hpp and libaether_hip.so, no HIP headers involved."""
    _host_driver("residency_check.cpp", RESIDENCY_CHECK)
    build_flag_check()
    return _host_driver("api_check.cpp", API_CHECK)


def _hip_driver(name: str, out: str) -> str:
    src = os.path.join(ROOT, "tests", "cpp", name)
    cmd = [HIPCC, *FLAGS, "-Werror", "-I", os.path.join(ROOT, "include"), src, "-o", out + ".tmp",
           "-L", LIB_DIR, "-laether_hip", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr}")
    os.replace(out + ".tmp", out)
    return out


def build_flag_check() -> str:
    """tests/cpp/flag_check.hip: a HIP program whose own kernels set aether::DeviceFlag through
    DeviceFlagView (include/aether/device/flag.hpp), the way a user's kernels would; and
    tests/cpp/expr_device_check.hip, whose own kernels evaluate the C++ front end's expressions per sample
    over device views (compared bit for bit with the whole-array assignment). hipcc against libaether_hip.so."""
    _hip_driver("expr_device_check.hip", EXPR_DEVICE_CHECK)
    return _hip_driver("flag_check.hip", FLAG_CHECK)


def build_dsc_core() -> str:
    """aether_amd.dsc._core (aether_amd/dsc/_core.cpp): the CPython extension of the array package, a
    host-only g++ build against Python's headers; it opens libaether_hip.so itself on the first device
    call, so it links neither the HIP runtime nor torch."""
    import sysconfig

    src = os.path.join(HERE, "dsc", "_core.cpp")
    out = os.path.join(HERE, "dsc", "_core" + sysconfig.get_config_var("EXT_SUFFIX"))
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", "-Wno-missing-field-initializers",
           "-Werror", "-fvisibility=hidden", "-I", sysconfig.get_paths()["include"], "-I", os.path.join(ROOT, "include"),
           src, "-o", out + ".tmp", "-ldl"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{res.stderr}")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
