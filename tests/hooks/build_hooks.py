"""Builds the user-kernel example the hook tests load (libtint_hook.so), with
the library's compiler flags and include/ on the include path."""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
LIB = HERE / "libtint_hook.so"


def build(hipcc: str, flags: list) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, *flags, "-shared", str(HERE / "tint_hook.hip"), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for tint_hook.hip:\n{res.stderr}")
    tmp.replace(LIB)
    return LIB
