"""The C-ABI library loads on a machine without a GPU and exports every entry point the
header declares, with the argument counts of the binding table (no compute calls)."""
import ctypes
import pathlib
import re

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared():
    names = {}
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\bint\s+(streams_\w+)\s*\(([^)]*)\)\s*;", text, flags=re.S):
            args = [a for a in m.group(2).split(",") if a.strip()]
            names[m.group(1)] = len(args)
    return names


def test_header_declares_entry_points():
    d = declared()
    assert "streams_eval_aux" in d and "streams_rosenbrock" in d
    assert len(d) >= 20


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.capi.load()
    for name in declared():
        assert hasattr(lib.lib, name), name


def test_binding_table_matches_header(pkg):
    d = declared()
    for name, kinds in pkg.capi.SIGNATURES.items():
        assert d["streams_" + name] == len(kinds), name


def test_oracle_exports_same_entry_points(pkg, oracle):
    """Every entry point that restates one reference routine has its CPU twin (the fused
    rhs_* stages are checked against sequences of those twins)."""
    for name in pkg.capi.SIGNATURES:
        if name.startswith("rhs_"):
            continue
        assert oracle.has(name), name
