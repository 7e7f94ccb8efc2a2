"""Build libjaqsi_amd.so (HIP kernels + C-ABI) in-tree for gfx950.

    python -m jaqsi_amd.build            # or __graft_entry__.build()

The library is written next to this file; the ctypes layer loads it from
there and raises if it is missing.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libjaqsi_amd.so")
BUILD = os.path.join(HERE, "_build")

SOURCES = ["plan.cpp", "capi.cpp", "jit.cpp", "sv_f32.hip", "sv_f64.hip", "prep.hip", "pulse.hip", "wide.hip", "evolve.hip", "noalloc.hip"]

# Headers the run-time compiler needs for plan-specialised kernels (jit.cpp):
# embedded in the library, so it does not depend on the source tree.
JIT_HEADERS = [("jaqsi_amd.h", INCLUDE), ("program.h", CSRC), ("sv_kernels.h", CSRC), ("sv_impl.h", CSRC)]

# Packed FP32 instructions stay off: hipcc -O3 would otherwise form them from
# plain float code.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the HIP toolchain is required to build jaqsi_amd")


def _compile(src: str) -> str:
    os.makedirs(BUILD, exist_ok=True)
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    path = os.path.join(CSRC, src)
    cmd = [hipcc(), *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c", path, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [hipcc(), "-O3", "-std=c++17", "-fPIC", "-I", INCLUDE, "-I", CSRC, "-I", BUILD, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def _embed_headers() -> None:
    """Write _build/jit_headers.inc: the kernel headers as raw string literals."""
    os.makedirs(BUILD, exist_ok=True)
    names, texts = [], []
    for name, d in JIT_HEADERS:
        text = open(os.path.join(d, name)).read()
        if ")JQHDR\"" in text:
            raise RuntimeError(f"{name} contains the raw-string delimiter")
        names.append(f'"{name}"')
        texts.append(f'R"JQHDR({text})JQHDR"')
    body = ("static const char* kJitHeaderNames[] = {" + ", ".join(names) + "};\n"
            "static const char* kJitHeaderTexts[] = {\n" + ",\n".join(texts) + "};\n")
    out = os.path.join(BUILD, "jit_headers.inc")
    if not os.path.exists(out) or open(out).read() != body:
        with open(out, "w") as fh:
            fh.write(body)


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    _embed_headers()
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
