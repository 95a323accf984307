"""Build libcudasph.so in-tree: HIP sources with hipcc for gfx950, host C++
with g++, linked by hipcc.  Run as ``python -m cudasph_amd.build``.

Every HIP translation unit is compiled with -ffp-contract=off (no FMA
contraction, so the plain form reproduces the reference's roundings; kernels
that want FMAs spell them out) and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
SRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libcudasph.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CUDASPH_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra", "-pthread"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    noise = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


MARKER = b"CUDASPH_SOURCE_HASH:"


def source_hash() -> str:
    """sha256 over the library's sources (csrc/ and include/), in name order."""
    h = hashlib.sha256()
    files = sorted(list(SRC.iterdir()) + list((PKG.parent / "include").glob("*.h")))
    for f in files:
        if f.suffix in (".hip", ".cpp", ".h"):
            h.update(f.name.encode() + b"\0" + f.read_bytes() + b"\0")
    return h.hexdigest()[:32]


def built_hash() -> str | None:
    """The source hash recorded in the built library, or None."""
    if not LIB.exists():
        return None
    data = LIB.read_bytes()
    k = data.find(MARKER)
    if k < 0:
        return None
    return data[k + len(MARKER): k + len(MARKER) + 32].decode("ascii", "replace")


def is_stale() -> bool:
    """True when the library is missing or was built from other sources."""
    return built_hash() != source_hash()


def ensure_built(verbose: bool = True) -> Path:
    """Build the library when it is missing or older than its sources."""
    if is_stale():
        return build(verbose)
    return LIB


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    stamp = OBJ / "source_hash.c"
    stamp.write_text('const char sph_source_hash_marker[] __attribute__((used)) = "'
                     + MARKER.decode() + source_hash() + '";\n')
    jobs = [["gcc", "-O2", "-fPIC", "-c", str(stamp), "-o", str(OBJ / "source_hash.c.o")]]
    for src in sorted(SRC.glob("*.hip")):
        jobs.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(OBJ / (src.stem + ".o"))])
    for src in sorted(SRC.glob("*.cpp")):
        jobs.append(["g++", *CXX_FLAGS, "-c", str(src), "-o", str(OBJ / (src.stem + ".cpp.o"))])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(_run, jobs))
    objs = [j[-1] for j in jobs]
    tmp = str(LIB) + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
