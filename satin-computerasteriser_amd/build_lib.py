"""Build libscr.so (HIP kernels + C ABI) for gfx950, in-tree.

Every translation unit is compiled by hipcc with FP contraction off (the
rasteriser's arithmetic contract) and without packed-FP32 instructions, then
linked into ``libscr.so`` next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
BUILD = PKG_DIR / "_build"
LIB = PKG_DIR / "libscr.so"
ARCH = os.environ.get("SCR_OFFLOAD_ARCH", "gfx950")

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{REPO / 'include'}",
    f"-I{CSRC}",
]


def source_hash() -> str:
    """First 12 hex digits of the SHA-256 of the library's sources (csrc and
    include, in name order): scr_version() reports it, so measurements kept
    beside the code can name the build they were taken on."""
    import hashlib
    h = hashlib.sha256()
    for f in sorted(list(CSRC.glob("*")) + list((REPO / "include").glob("*.h"))):
        if f.is_file() and f.suffix in (".hip", ".hpp", ".cpp", ".h"):
            h.update(f.name.encode())
            h.update(f.read_bytes())
    return h.hexdigest()[:12]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


# scr_mv_draw.hip holds the batched draw kernel of one (phase, shape) pair
# and is compiled once per pair: 3 phases x 6 shapes
MV_DRAW = "scr_mv_draw.hip"
MV_UNITS = [(ph, sh) for ph in range(3) for sh in range(6)]


def _compile(unit, hipcc: str, src_hash: str) -> Path:
    src, defines, stem = unit
    obj = BUILD / (stem + ".o")
    cmd = [hipcc, *COMMON_FLAGS, f'-DSCR_SOURCE_HASH="{src_hash}"', *defines, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {stem}:\n{res.stderr}")
    return obj


def _units():
    units = []
    # *.hip: kernels and their host code; *.cpp: host-only code (PLY input)
    for src in sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp"))):
        if src.name == MV_DRAW:
            units += [(src, [f"-DSCR_MV_PHASE={ph}", f"-DSCR_MV_SHAPE={sh}"], f"{src.stem}_{ph}_{sh}")
                      for ph, sh in MV_UNITS]
        else:
            units.append((src, [], src.stem))
    # longest first: the multi-view draw units, then the rest
    return sorted(units, key=lambda u: 0 if u[0].name == MV_DRAW else 1)


def build(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    units = _units()
    with cf.ThreadPoolExecutor(max_workers=min(os.cpu_count() or 8, 16, len(units))) as ex:
        src_hash = source_hash()
        objs = list(ex.map(lambda u: _compile(u, hipcc, src_hash), units))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(units)} objects", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
