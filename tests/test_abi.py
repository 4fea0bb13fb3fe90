"""The C-ABI library loads on the CPU host and exports every symbol that
include/*.h declares (no compute calls without a GPU)."""

import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names.update(re.findall(r"\b(cia_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from concrete_impact_ai_amd import _lib

    handle = _lib.lib()
    missing = [n for n in declared_symbols() if not hasattr(handle, n)]
    assert not missing, missing
    assert handle.cia_abi_version() == 1


def test_bindings_cover_declarations():
    from concrete_impact_ai_amd import _lib

    _lib.lib()
    assert set(declared_symbols()) <= set(_lib.exported_symbols())


def test_struct_layouts_match_header():
    """Compile a C probe against include/cia.h and compare sizeof with ctypes."""
    import subprocess
    import tempfile

    from concrete_impact_ai_amd import _lib

    src = r"""
#include <stdio.h>
#include "cia.h"
int main(){printf("%zu %zu %zu %zu %zu %zu\n", sizeof(cia_rve_result), sizeof(cia_rve_desc),
 sizeof(cia_point_failure), sizeof(cia_material), sizeof(cia_update_settings),
 sizeof(cia_surrogate_gate));}
"""
    with tempfile.TemporaryDirectory() as d:
        c = Path(d) / "p.c"
        c.write_text(src)
        exe = Path(d) / "p"
        subprocess.run(["gcc", "-I", str(ROOT / "include"), str(c), "-o", str(exe)], check=True)
        sizes = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                                check=True).stdout.split()]
    assert sizes == [ctypes.sizeof(_lib.RVEResult), ctypes.sizeof(_lib.RVEDesc),
                     ctypes.sizeof(_lib.PointFailure), ctypes.sizeof(_lib.Material),
                     ctypes.sizeof(_lib.UpdateSettings), ctypes.sizeof(_lib.SurrogateGate)]
