"""Builds libezlda.so (HIP kernels + C ABI, gfx950) and the `ezlda` CLI in-tree.

    python -m lda_sc20_amd.build        # or lda_sc20_amd.build.build()

Float rules shared with the CPU oracle: no FMA contraction
(-ffp-contract=off), correctly rounded f32 division and square root.  Packed
FP32 instructions are disabled for device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
INCLUDE = os.path.join(os.path.dirname(PKG), "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libezlda.so")
CLI = os.path.join(PKG, "ezlda")
ARCH = os.environ.get("EZLDA_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          f"-I{INCLUDE}", f"-I{CSRC}"]
DEVICE = [f"--offload-arch={ARCH}", "-fhip-fp32-correctly-rounded-divide-sqrt",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          ]

SOURCES = [
    ("ezlda_kernels.hip", True),
    ("ezlda_runtime.cpp", True),
    ("ezlda_corpus.cpp", False),
    ("ezlda_foldin.cpp", True),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src, device):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON
    if device:
        cmd += DEVICE + ["-x", "hip"]
    else:
        cmd += ["-x", "c++"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    _run(cmd)
    return obj


def build(verbose: bool = False) -> str:
    """Compiles every source and links libezlda.so and the ezlda CLI."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = [(s, d) for s, d in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=len(srcs)) as ex:
        objs = list(ex.map(lambda sd: _compile(*sd), srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    cli_src = os.path.join(HOST, "ezlda_cli.cpp")
    if os.path.exists(cli_src):
        tmpc = CLI + ".tmp"
        _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", f"-I{HOST}", cli_src, "-o", tmpc,
              f"-L{PKG}", "-lezlda", "-Wl,-rpath,$ORIGIN", "-lpthread"])
        os.replace(tmpc, CLI)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
