"""The C-ABI library loads without a GPU and exports every symbol include/*.h declares."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        # pf_* entry points and the exports under the reference's own C names
        for m in re.finditer(r"\b((?:pf_|crs_spmv_)[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "pf_bicgstab_crs" in names and "pf_crs_graph" in names and len(names) >= 25
    assert {"crs_spmv_real64", "crs_spmv_real32", "pf_crs_spmv_real64", "pf_crs_spmv_real32"} <= set(names)


def test_library_exports_every_declared_symbol():
    from plantfem_amd import _lib as L
    lib = L.load()  # dlopen only: no device call is made
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from plantfem_amd import _lib as L
    assert set(declared_functions()) <= set(L.SIGNATURES)


def test_options_defaults_without_gpu():
    from plantfem_amd import _lib as L
    o = L.default_opts()
    assert o.itrmax == 100000 and o.precision == 64
    # dble(1.0e-10): the REAL(4) literal widened (FEMSolverClass.f90:70-71)
    assert o.er0 == o.relative_er == float.fromhex("0x1.b7cdfe0000000p-34")
    assert ctypes.sizeof(L.SolveOpts) == 4 * 2 + 8 * 4 + 4 * 10


def test_no_oracle_in_product():
    """The product package never imports, links or loads the oracle."""
    for p in (ROOT / "plantfem_amd").rglob("*"):
        if p.suffix in (".py", ".hip", ".hpp", ".cpp", ".f90", ".h"):
            assert "oracle" not in p.read_text().replace("oracle is", ""), p
