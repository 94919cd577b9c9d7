"""The C-ABI library loads without a GPU, exports every entry point include/firepro.h
declares, and the ctypes mirror matches the header's struct layout."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from firepro import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "firepro.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(fp_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(abi.LIBRARY):
        pytest.skip("libfirepro.so is not built")
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIBRARY], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (fp_[a-z_]+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    lib = abi.load()  # loads on a CPU-only host: no device call happens here
    lib.fp_version.restype = C.c_char_p
    assert b"gfx950" in lib.fp_version()


@pytest.mark.parametrize("name,ctype", [
    ("fp_grid", abi.Grid), ("fp_settings", abi.Settings), ("fp_flame", abi.Flame),
    ("fp_fuel", abi.Fuel), ("fp_source", abi.Source), ("fp_force", abi.Force),
    ("fp_collider", abi.Collider), ("fp_frame", abi.Frame), ("fp_stats", abi.Stats),
    ("fp_tables", abi.Tables), ("fp_camera", abi.Camera), ("fp_render_settings", abi.RenderSettings),
    ("fp_render_bricks", abi.RenderBricks), ("fp_render_tables", abi.RenderTables),
])
def test_struct_layout_matches_header(name, ctype):
    src = f'#include "{HEADER}"\n#include <stdio.h>\nint main(void){{printf("%zu", sizeof({name}));return 0;}}\n'
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "s")
        subprocess.run(["gcc", c, "-o", exe], check=True)
        size = int(subprocess.run([exe], capture_output=True, text=True, check=True).stdout)
    assert C.sizeof(ctype) == size


def test_library_fails_loudly_when_unbuilt(tmp_path):
    import importlib
    mod = importlib.reload(abi)
    saved = mod._lib
    mod._lib = None
    try:
        with pytest.raises(mod.FireProError, match="missing"):
            mod.load(str(tmp_path / "libfirepro.so"))
    finally:
        mod._lib = saved
