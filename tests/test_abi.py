"""The C-ABI library loads and exports every symbol include/m1march.h declares.

No compute call is made: these run without a GPU.
"""

import ctypes
import re
from pathlib import Path

import pytest

from helpers import ROOT

HEADER = ROOT / "include" / "m1march.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(m1_[a-z_]+)\s*\(", text)))


def test_header_declares_the_march_entry_points():
    names = declared_functions()
    for required in ("m1_create", "m1_set_problem", "m1_march", "m1_spatial_operator",
                     "m1_compute_energy_step", "m1_step", "m1_get_dose", "m1_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from proton_solver import _native

    lib = _native.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.m1_abi_version() == _native.ABI_VERSION


def test_python_binding_covers_the_header():
    from proton_solver import _native

    bound = {name for name, _, _ in _native.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_struct_layouts_match_header():
    """ctypes structures have the C layout of the header structs."""
    from proton_solver import _native

    assert ctypes.sizeof(_native.m1_table) == 32
    assert ctypes.sizeof(_native.m1_material) == 3 * 32 + 16
    assert ctypes.sizeof(_native.m1_beam) == 24
    assert ctypes.sizeof(_native.m1_numerics) == 6 * 4 + 5 * 8 + 2 * 4
    assert _native.m1_numerics.max_energy_step.offset == 24
    assert _native.m1_numerics.vacuum_aware_cfl.offset == 64
    assert ctypes.sizeof(_native.m1_beamlet) == 6 * 4 + 24
    assert _native.m1_beamlet.beam.offset == 24


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """sizeof / offsetof of the header structs, as gcc lays them out, equal
    the ctypes layouts."""
    import shutil
    import subprocess

    from helpers import ROOT
    from proton_solver import _native

    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    structs = {"m1_table": _native.m1_table, "m1_material": _native.m1_material,
               "m1_beam": _native.m1_beam, "m1_numerics": _native.m1_numerics,
               "m1_beamlet": _native.m1_beamlet}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "m1march.h"', "int main(void) {"]
    for name, cls in structs.items():
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{name}.{fname} %zu\\n", offsetof({name}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = dict(ln.split() for ln in subprocess.run([str(exe)], check=True, capture_output=True,
                                                    text=True).stdout.splitlines())
    for name, cls in structs.items():
        assert int(got[name]) == ctypes.sizeof(cls), name
        for fname, _ in cls._fields_:
            assert int(got[f"{name}.{fname}"]) == getattr(cls, fname).offset, (name, fname)


def test_library_kernel_names_without_gpu():
    from proton_solver import _native

    lib = _native.load_library()
    names = [lib.m1_kernel_name(k).decode() for k in range(lib.m1_kernel_count())]
    assert names[:3] == ["stage0", "stage1", "stage2"]
