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
    rhs_* stages are checked against sequences of those twins, the reactor batch against the
    oracle's per-reactor Rosenbrock, the allocation-free variants against their allocating
    entry points)."""
    for name in pkg.capi.SIGNATURES:
        if name.startswith("rhs_") or name == "reactor_batch" or name.endswith("_noalloc"):
            continue
        assert oracle.has(name), name


def test_kinetics_pack_restates_the_dense_tables(pkg):
    """streams_kinetics_pack is host code: its records decode back to the reference's dense
    kinetics tables (stoichiometry, net coefficients times W_s, rates, efficiencies)."""
    import numpy as np
    for mech in (pkg.cases.h2_air_mechanism(),
                 pkg.mechanism.Mechanism(str(pkg.cases.mechanism_path("keromnes2013_14sp.yaml")))):
        tb = mech.tables(transport=False)
        ns, nr = tb.n_species, tb.nreactions
        pack = pkg.capi.kinetics_pack(tb).reshape(nr, 80)
        ints = pack[:, :16].copy().view(np.int32)
        r_co, p_co = np.zeros((nr, ns)), np.zeros((nr, ns))
        for r in range(nr):
            assert ints[r, 0] == tb.reac_ty[r] and ints[r, 1] == tb.isRev[r]
            for e in range(ints[r, 2]):
                r_co[r, ints[r, 8 + e]] = pack[r, 28 + e]
            for e in range(ints[r, 3]):
                p_co[r, ints[r, 12 + e]] = pack[r, 32 + e]
            np.testing.assert_array_equal(pack[r, 36:36 + ns], (tb.p_coeffs[r] - tb.r_coeffs[r]) * tb.mw)
            np.testing.assert_array_equal(pack[r, 16:19], [tb.arr_a[r, 0], tb.arr_b[r, 0], tb.arr_ea[r, 0]])
            np.testing.assert_array_equal(pack[r, 19:22], [tb.arr_a[r, 1], tb.arr_b[r, 1], tb.arr_ea[r, 1]])
            np.testing.assert_array_equal(pack[r, 22:27], tb.falloff_coeffs[r])
            np.testing.assert_array_equal(pack[r, 52:52 + ns], tb.tb_eff[r])
            assert ints[r, 5] == int(tb.arr_b[r, 0] != 0 or tb.arr_ea[r, 0] != 0)
            assert ints[r, 6] == int(tb.falloff_coeffs[r, 3] != -3.14)
        np.testing.assert_array_equal(r_co, tb.r_coeffs)
        np.testing.assert_array_equal(p_co, tb.p_coeffs)


def test_bench_refuses_more_than_one_gpu():
    """bench.py drives one GPU; --gpus 2 exits with status 2 and says so, before touching a GPU."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "--gpus 2 is not supported" in r.stderr
    assert r.stdout == ""
