"""Headless runs of the reference's presets (examples/playback/main.ts over
tests/golden/presets): the HIP path against the FP32 restatement, bit-exact."""
import glob
import json
import os

import pytest

from lockstep import run_lockstep

pytestmark = pytest.mark.gpu

PRESETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "presets")
TEXT_SHAPED = {"flaming-letters"}  # its emitter needs the editor's font geometry
NAMES = sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(PRESETS, "*.json")))


@pytest.mark.parametrize("name", [n for n in NAMES if n not in TEXT_SHAPED])
def test_preset_bit_exact(name):
    from firepro import document
    with open(os.path.join(PRESETS, name + ".json")) as f:
        doc = json.load(f)
    print(f"case preset {name}", flush=True)
    assert run_lockstep(document.preset_scene(doc, 12), check_every=4) == []
