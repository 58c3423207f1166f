"""Builds the user-kernel examples the hook tests load (libuser_hooks.so: a
tint kernel and a displacement kernel), with the library's compiler flags and
include/ on the include path."""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
LIB = HERE / "libuser_hooks.so"
SOURCES = ("tint_hook.hip", "displace_hook.hip")


def build(hipcc: str, flags: list) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, *flags, "-shared", *[str(HERE / s) for s in SOURCES], "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for the user hooks:\n{res.stderr}")
    tmp.replace(LIB)
    return LIB
