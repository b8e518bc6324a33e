"""fluid_prop.bin: the Fortran sequential-unformatted table file that set_fluid_prop reads
(multideal.F90:2763-2926) and tools/preproc_multi/preproc.py:196-246 writes.  The record
sequence and every record's byte length must be what the Fortran reader's declarations
take, or the reference reads shifted data."""
import struct

import numpy as np
import pytest


def _records(raw):
    out, off = [], 0
    while off < len(raw):
        (n,) = struct.unpack_from("<i", raw, off)
        (m,) = struct.unpack_from("<i", raw, off + 4 + n)
        assert m == n, "leading and trailing record markers differ"
        out.append(raw[off + 4: off + 4 + n])
        off += 8 + n
    return out


@pytest.fixture(scope="module")
def written(pkg, tmp_path_factory):
    tb = pkg.cases.h2_air_mechanism().tables(t_max_tab=3000.0)
    path = tmp_path_factory.mktemp("fp") / "fluid_prop.bin"
    tb.write_fluid_prop(str(path))
    return tb, path


def test_record_layout_matches_the_reader(written):
    tb, path = written
    ns, nr, n1 = tb.n_species, tb.nreactions, tb.num_t_tab + 1
    ncoef = tb.indx_cp_r + 2 - tb.indx_cp_l + 1
    want = [
        8,                       # read(10) number_species, nreactions        (2 x int32)
        30 * ns,                 # species_names: character(len=30)(N_S)
        8 * ns, 8 * ns, 8 * ns,  # mw, rgas, h298
        8 * ns, 8 * ns, 8 * ns, 8 * ns,   # sigma, epsK, dipole, polariz
        4 * ns,                  # geometry (integer)
        8,                       # number_cp_coeff, nsetcv
        8 * ns * (tb.nsetcv + 1),           # trange(N_S, nsetcv+1)
        8 * ncoef * ns * tb.nsetcv,         # cp_coeff(indx_cp_l:indx_cp_r+2, N_S, nsetcv)
        24,                      # t_min_tab, t_max_tab, dt_tab
        8 * n1 * ns, 8 * n1 * ns, 8 * n1 * ns * ns,   # visc, lambda, diffbin
        4 * nr,                  # reac_ty
        16 * nr, 16 * nr, 16 * nr,          # arr_a, arr_b, arr_ea (nr, 2)
        40 * nr,                 # falloff_coeffs(nr, 5)
        8 * nr * ns, 8 * nr * ns, 8 * nr * ns,   # tb_eff, r_coeffs, p_coeffs
        8 * n1 * nr,             # kc_tab(num_t_tab+1, nr)
    ]
    recs = _records(path.read_bytes())
    assert [len(r) for r in recs] == want
    assert tuple(np.frombuffer(recs[0], dtype=np.int32)) == (ns, nr)
    names = [recs[1][30 * i:30 * (i + 1)].decode() for i in range(ns)]
    assert [n.rstrip() for n in names] == list(tb.names)
    assert all(len(n) == 30 for n in names)
    # cp coefficients are stored divided by the species gas constant (the reader
    # multiplies them back, multideal.F90:2806-2811)
    cp = np.frombuffer(recs[12], dtype=np.float64).reshape((ncoef, ns, tb.nsetcv), order="F")
    np.testing.assert_allclose(cp * tb.rgas[None, :, None], tb.cp_coeff, rtol=1e-15)
    # transport data: H2's Lennard-Jones diameter 2.92 Angstrom, well depth 38 K
    sig = np.frombuffer(recs[5], dtype=np.float64)
    k = [n.upper() for n in tb.names].index("H2")
    assert sig[k] == pytest.approx(2.92e-10) and np.frombuffer(recs[6])[k] == pytest.approx(38.0)


def test_round_trip(pkg, written):
    tb, path = written
    back = pkg.mechanism.read_fluid_prop(str(path), tb.p0)
    assert back.names == tb.names
    # binary diffusivities go through the 1 atm scaling and back (two roundings)
    np.testing.assert_allclose(back.diffbin_species, tb.diffbin_species, rtol=4e-16, atol=0)
    for key in ("mw", "rgas", "h298", "trange", "cp_coeff", "cv_coeff", "visc_species", "lambda_species",
                "arr_a", "arr_b", "arr_ea", "falloff_coeffs", "tb_eff", "r_coeffs",
                "p_coeffs", "sigma", "eps_k", "dipole", "polariz", "geometry"):
        np.testing.assert_array_equal(getattr(back, key), getattr(tb, key), err_msg=key)
    np.testing.assert_array_equal(back.reac_ty, tb.reac_ty)
    # the reader clamps K_c to [tiny, huge] (multideal.F90:2918-2923)
    np.testing.assert_array_equal(back.kc_tab, np.clip(tb.kc_tab, np.finfo(float).tiny, np.finfo(float).max))
    assert (back.num_t_tab, back.dt_tab, back.t_min_tab) == (tb.num_t_tab, tb.dt_tab, tb.t_min_tab)


def test_reader_rejects_other_name_widths(pkg, written, tmp_path):
    tb, path = written
    raw = bytearray(path.read_bytes())
    recs = _records(bytes(raw))
    ns = tb.n_species
    bad = b"".join(n.ljust(40).encode() for n in tb.names)
    out = bytearray()
    for i, r in enumerate(recs):
        r = bad if i == 1 else r
        out += struct.pack("<i", len(r)) + r + struct.pack("<i", len(r))
    p = tmp_path / "fp40.bin"
    p.write_bytes(bytes(out))
    with pytest.raises(ValueError, match=f"{ns} x 30"):
        pkg.mechanism.read_fluid_prop(str(p), tb.p0)
