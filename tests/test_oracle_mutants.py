"""The oracle's pins have teeth: the oracle rebuilt with each planted mistake
(scr_oracle.c under -DORC_MUTANT=k) must fail the pin that covers it.  CPU
only; the planted mistakes never reach the default oracle build."""
import shutil
import subprocess
from pathlib import Path

import pytest

import test_oracle_pinning as pins
from _pkg import orc

REPO = Path(__file__).resolve().parents[1]
CC = shutil.which("gcc") or shutil.which("cc")

# mutant -> (what it plants, the pin that must fail)
MUTANTS = {
    1: ("a frustum plane built from the wrong clip row", pins.pin_frustum),
    2: ("floor instead of ceil in the footprint radius", pins.pin_footprint_radius),
    3: ("nearest ties resolved to the highest index", pins.pin_nearest_matches_cpu_reference),
    4: ("the pixel size's x extent scaled by the height", pins.pin_pixel_size_and_precision),
    5: ("the CLOD threshold without its + 3", pins.pin_lod_threshold),
    6: ("blend sums not split over the sweep's samples", pins.pin_blend_sums_and_coverage),
    7: ("coverage alpha from the fixed-point sum unscaled", pins.pin_blend_sums_and_coverage),
    8: ("precision thresholds 100 and 200 swapped", pins.pin_pixel_size_and_precision),
    9: ("the motion-blur sweep in the wrong direction", pins.pin_blend_sums_and_coverage),
    10: ("world-space point size from the view distance instead of the depth", pins.pin_footprint_radius),
    11: ("the near plane in the OpenGL w + z form", pins.pin_frustum),
    12: ("no NaN cull of displaced points", pins.pin_displacement),
    13: ("the motion-blur sweep from the displaced point instead of position + previous displacement",
         pins.pin_displacement),
    19: ("the batch projection's origin without the transform's translation", pins.pin_batch_projection),
    20: ("the batch projection's y column scaled by the x step", pins.pin_batch_projection),
}

# planted in the composite's numpy restatement (oracle.COMPOSITE_MUTANT = k)
PY_MUTANTS = {
    14: ("source-over's destination weight a instead of 1 - a", pins.pin_composite),
    15: ("source-over alpha without the destination term", pins.pin_composite),
    16: ("depth 0 texels drawn in the depth mode", pins.pin_composite),
    17: ("the greater compare passing ties", pins.pin_composite),
    18: ("the depth written where the compare fails", pins.pin_composite),
}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    if CC is None:
        pytest.skip("no C compiler")
    out = tmp_path_factory.mktemp("oracle_mutants")
    libs = {}
    for k in MUTANTS:
        so = out / f"libscr_oracle_m{k}.so"
        subprocess.run([CC, "-O1", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-D_GNU_SOURCE",
                        f"-DORC_MUTANT={k}", "-shared", "-o", str(so), str(REPO / "oracle" / "scr_oracle.c"),
                        "-lm", "-lpthread"], check=True, capture_output=True)
        libs[k] = so
    return libs


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_each_planted_mistake_fails_its_pin(mutant_libs, k):
    what, pin = MUTANTS[k]
    prev = orc.use_library(mutant_libs[k])
    try:
        with pytest.raises(AssertionError):
            pin()
    finally:
        orc.use_library(prev)
    pin()  # and the real oracle passes it


@pytest.mark.parametrize("k", sorted(PY_MUTANTS))
def test_each_planted_composite_mistake_fails_its_pin(k):
    what, pin = PY_MUTANTS[k]
    orc.COMPOSITE_MUTANT = k
    try:
        with pytest.raises(AssertionError):
            pin()
    finally:
        orc.COMPOSITE_MUTANT = 0
    pin()
