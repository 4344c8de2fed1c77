"""The C ABI of include/sphmhd.h: the built library exports every declared
entry point, and the Python (ctypes) view agrees with the header's enums and
struct layouts (checked against offsetof from the C compiler).  No compute
call is made here (these run without a GPU)."""
from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "sphmhd.h"
LIB = ROOT / "pysph-mhd_amd" / "sphmhd" / "libsphmhd.so"


def _declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sphmhd_\w+)\s*\(", text, re.M)))


def test_header_declares_the_entry_points():
    names = _declared_functions()
    for must in ("sphmhd_create", "sphmhd_bind", "sphmhd_eval", "sphmhd_step", "sphmhd_warm_start",
                 "sphmhd_pec_stage1", "sphmhd_pec_stage2", "sphmhd_neighbours", "sphmhd_cells",
                 "sphmhd_last_error", "sphmhd_totals"):
        assert must in names


@pytest.mark.skipif(not LIB.exists(), reason="libsphmhd.so not built (run __graft_entry__.build())")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [n for n in _declared_functions() if n not in exported]
    assert not missing, missing


@pytest.mark.skipif(not LIB.exists(), reason="libsphmhd.so not built")
def test_library_loads_and_reports_abi_version():
    lib = ctypes.CDLL(str(LIB))
    lib.sphmhd_abi_version.restype = ctypes.c_int
    from sphmhd import _abi

    assert lib.sphmhd_abi_version() == _abi.ABI_VERSION
    for name in _declared_functions():
        assert hasattr(lib, name), name


def test_field_enum_matches_python_names():
    from sphmhd import _abi

    text = HEADER.read_text()
    body = text[text.index("typedef enum sphmhd_field"): text.index("} sphmhd_field;")]
    names = re.findall(r"SPHMHD_F_(\w+)", body)
    assert [n.lower() for n in names] == [f.lower() for f in _abi.FIELDS]
    assert "SPHMHD_NFIELDS" in body and _abi.NFIELDS == len(names)


def test_totals_enum_matches():
    from sphmhd import _abi

    text = HEADER.read_text()
    body = text[text.index("SPHMHD_T_MASS"): text.index("SPHMHD_NTOTALS")]
    names = re.findall(r"SPHMHD_T_(\w+)", body)
    assert [n.lower() for n in names] == list(_abi.TOTALS)


def _c_offsets(tmp_path, struct, members):
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {",
             f'  printf("%zu\\n", sizeof({struct}));']
    for m in members:
        lines.append(f'  printf("%zu\\n", offsetof({struct}, {m}));')
    lines += ["  return 0;", "}"]
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    return vals[0], vals[1:]


@pytest.mark.parametrize("cname,pyname", [("sphmhd_params", "Params"), ("sphmhd_stats", "Stats"),
                                          ("sphmhd_timing", "Timing")])
def test_struct_layout_matches_header(tmp_path, cname, pyname):
    from sphmhd import _abi

    cls = getattr(_abi, pyname)
    members = [m for m, _ in cls._fields_]
    size, offs = _c_offsets(tmp_path, cname, members)
    assert ctypes.sizeof(cls) == size
    assert [getattr(cls, m).offset for m in members] == offs


def test_oracle_uses_the_same_parameter_layout(tmp_path):
    """The oracle includes include/sphmhd.h for sphmhd_params, so one ctypes
    Structure feeds both."""
    text = (ROOT / "oracle" / "sphmhd_oracle.h").read_text()
    assert '#include "../include/sphmhd.h"' in text
