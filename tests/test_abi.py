"""The C ABI library loads and exports every function include/scr.h declares
(CPU only: no compute calls)."""
import ctypes
import re
from pathlib import Path

from _pkg import scr

REPO = Path(__file__).resolve().parents[1]


def _declared_functions():
    text = (REPO / "include" / "scr.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(scr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_core_entry_points():
    names = _declared_functions()
    for must in ("scr_ctx_create", "scr_cloud_add_batches", "scr_render", "scr_pick", "scr_pack", "scr_frame_read"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(scr.LIB_PATH))
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_every_declared_symbol():
    lib = scr.load_library()
    assert set(_declared_functions()) <= set(scr._SIGNATURES)
    assert lib.scr_version().decode().startswith("scr")


def test_config_defaults_match_reference():
    c = scr.CConfig()
    scr.load_library().scr_config_default(ctypes.byref(c))
    py = scr.ComputeRasteriserConfiguration().to_c()
    for name, _ in scr.CConfig._fields_:
        if name in ("backgroundColor", "reserved"):
            continue
        assert getattr(c, name) == getattr(py, name), name
    # ComputeRasteriserTypes.swift:97-119
    assert abs(c.depthTolerance - 0.01) < 1e-9 and c.enableFrustumCulling == 1 and c.enableCLOD == 1
    assert c.minimumPointSize == 1.0 and c.maximumPointSize == 1.0 and c.pointSizeScale == 1.0
    assert c.holeFillIterations == 0 and c.motionBlurSamples == 8 and c.motionBlurMaxSpread == 64.0
    assert ctypes.sizeof(scr.CConfig) == 96
