"""Builds the library's sources against the CPU emulation of HIP in this directory, with
AddressSanitizer and UndefinedBehaviorSanitizer, into one test executable (TEST
INFRASTRUCTURE; never shipped, never loaded by the product).

The only source transformation is mechanical: `extern __shared__ T name[];` (dynamic LDS)
becomes a pointer to the launch's emulated LDS block, since a C++ host compiler has no
address space for it.
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "msplat_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
OUT = os.environ.get("MSPLAT_EMU_OUT", os.path.join(HERE, "_build"))
EXE = os.path.join(OUT, "emu_driver")
SHLIB = os.path.join(OUT, "libmsplat_emu.so")
SOURCES = ["kernels.hip", "mcmc.hip", "mlp.hip", "eval.hip", "render_noalloc.hip", "engine.cpp", "batch.cpp", "proposal.cpp", "model.cpp", "api.cpp", "scene.cpp", "io.cpp", "image_io.cpp", "ingest.cpp", "io_api.cpp"]
FLAGS = ["-std=c++20", "-O1", "-g", "-fPIC", "-U_FORTIFY_SOURCE", "-D_FORTIFY_SOURCE=0", "-fno-omit-frame-pointer", "-fsanitize=address",
         "-ffp-contract=off", "-pthread", "-Wno-deprecated-declarations", "-I" + HERE, "-I" + INCLUDE, "-I" + CSRC]

_DYN = re.compile(r"extern\s+__shared__\s+([A-Za-z_][A-Za-z0-9_]*)\s+([A-Za-z_][A-Za-z0-9_]*)\s*\[\s*\]\s*;")


def _prepare(src: str) -> str:
    text = open(os.path.join(CSRC, src)).read()
    text = _DYN.sub(lambda m: f"{m.group(1)}* {m.group(2)} = ({m.group(1)}*)emu_dyn_smem();", text)
    dst = os.path.join(OUT, src + ".cpp")
    if not os.path.exists(dst) or open(dst).read() != text:
        with open(dst, "w") as f:
            f.write(text)
    return dst


def _compile(path: str) -> str:
    obj = os.path.join(OUT, os.path.basename(path) + ".o")
    if os.path.exists(obj) and os.path.getmtime(obj) >= os.path.getmtime(path) and path.endswith(".hip.cpp"):
        deps = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
        deps.append(os.path.join(HERE, "hip", "hip_runtime.h"))
        if all(os.path.getmtime(obj) >= os.path.getmtime(d) for d in deps):
            return obj
    cmd = ["g++"] + FLAGS + ["-x", "c++", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"emulation build failed for {path}:\n{r.stderr[-6000:]}")
    return obj


def build() -> str:
    os.makedirs(OUT, exist_ok=True)
    paths = []
    for s in SOURCES:
        paths.append(_prepare(s) if s.endswith(".hip") else os.path.join(CSRC, s))
    paths.append(os.path.join(HERE, "emu_driver.cpp"))
    with ThreadPoolExecutor(max_workers=min(8, len(paths))) as ex:
        objs = list(ex.map(_compile, paths))
    cmd = ["g++"] + FLAGS + objs + ["-o", EXE, "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"emulation link failed:\n{r.stderr[-6000:]}")
    cmd = ["g++"] + FLAGS + ["-shared"] + objs[:-1] + ["-o", SHLIB, "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"emulation shared-library link failed:\n{r.stderr[-6000:]}")
    return EXE


def asan_runtime() -> str:
    """Path of the AddressSanitizer runtime, to LD_PRELOAD into a Python that loads SHLIB."""
    return subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()


if __name__ == "__main__":
    print(build())
    sys.exit(0)
