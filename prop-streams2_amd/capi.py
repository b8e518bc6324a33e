"""ctypes binding of ``libstreams_hip.so`` (the C-ABI declared in ``include/streams_hip.h``).

``SIGNATURES`` lists every entry point with its argument kinds in header order; the same
table drives the binding here and any other library that exports the same entry points
under another prefix.  Argument kinds: ``i`` int, ``d`` double, ``p`` pointer (torch
tensor, numpy array, integer address or None), ``s`` stream, ``P`` host pointer to an
output array.

Loading fails loudly: there is no fallback path when the library is missing or a GPU is
absent.
"""
from __future__ import annotations

import ctypes
import os
import pathlib
from typing import Dict, List, Tuple

PKG = pathlib.Path(__file__).resolve().parent
LIBPATH = PKG / "lib" / "libstreams_hip.so"

STATUS = {-1: "N_S not compiled", -2: "inconsistent sizes", -3: "unsupported option", -4: "allocation failed"}

_ARGS = {
    "eval_aux": "iiiiiiiiiiiii" "pp" "dddd" "ppp" "ip" "iid" "s" "ppppp" "i" "i" "pppp",
    "eval_aux2": "iiiiiii" "p" "pppp" "d" "i" "s",
    "eval_velaux": "iiiiiii" "pp" "s",
    "les_wale_mut": "iiiiiii" "p" "pppp" "dd" "i" "s",
    "eval_aux_les": "iiiiiii" "pp" "dddd" "ppp" "i" "p" "ii" "d" "ppppp" "i" "dd" "s",
    "eval_chem_aux": "iiiiiii" "p" "p" "ppp" "pp" "pp" "ppp" "i" "ddd" "ppp" "p" "i" "ddd" "s",
    "bcextr_var": "iiii" "p" "s",
    "bcswap": "iiiii" "p" "p" "s",
    "bcextr": "iiiiii" "p" "s",
    "bcwall_ad": "iiiiii" "p" "s",
    "bc_nr_lat": "iiiiiiiii" "ppp" "p" "ii" "ppp" "i" "pp" "s",
    "euler_x": "iiiiiii" "iiii" "pppp" "p" "ii" "ii" "d" "i" "pp" "i" "p" "ii" "p" "ddd" "i" "ii" "p" "s",
    "euler_y": "iiiiiii" "iiii" "pp" "p" "pp" "p" "ii" "ii" "d" "i" "pp" "i" "p" "ii" "p" "ddd" "i" "s",
    "euler_z": "iiiiiii" "iiii" "pp" "p" "pp" "p" "ii" "ii" "d" "i" "pp" "i" "p" "ii" "p" "ddd" "i" "s",
    "visflx_x": "iiiiiii" "p" "p" "pp" "ii" "p" "i" "pp" "d" "s",
    "visflx_y": "iiiiiii" "p" "p" "pp" "ii" "p" "i" "pp" "d" "s",
    "visflx_z": "iiiiiii" "p" "p" "pp" "ii" "p" "i" "pp" "d" "s",
    "visflx_reduced": "iiiiiii" "p" "p" "pppp" "s",
    "visflx": "iiiiiii" "pp" "pp" "ppppppppp" "s",
    "compute_chemistry": "iiiiiiii" "ppp" "p" "pppp" "p" "pp" "ppp" "i" "ddd" "ii" "p" "i" "p" "d" "s",
    "rosenbrock": "iiiiiii" "ppp" "p" "i" "pppp" "p" "pp" "ppp" "i" "dddd" "p" "i" "p" "ii" "ddd" "iid" "ii"
                  "ddd" "ppp" "s",
    "rosenbrock_krylov": "iiiiiii" "ppp" "p" "i" "pppp" "p" "pp" "ppp" "i" "dddd" "p" "i" "p" "ii" "ddd" "i" "p"
                         "ii" "ddd" "ppp" "s",
    "init_flux": "iiii" "pp" "d" "s",
    "update_flux": "iiii" "pp" "d" "s",
    "init_flux_simpler": "iiii" "ppp" "d" "s",
    "sav_flx": "iiii" "pp" "s",
    "update_simpler_flux": "iiii" "ppp" "d" "s",
    "update_field": "iiiii" "ppp" "s",
    "compute_dt": "iiiiiii" "pppppp" "pp" "P" "ii" "p" "i" "pp" "s",
    "compute_dt_noalloc": "iiiiiii" "pppppp" "pp" "P" "ii" "p" "i" "pp" "s",
    "copy_cpu_gpu": "iiiii" "P" "p" "s",
    "copy_gpu_cpu": "iiiii" "p" "P" "s",
    "field_interior": "iiiii" "pp" "i" "s",
    "compute_residual": "iiiii" "p" "d" "P" "p" "s",
    "compute_rho_t_p_minmax": "iiiiii" "p" "PPPPPP" "p" "s",
    "find_boom": "iiiii" "p" "d" "p" "ii" "PP" "s",
    "rhs_sweep": "iiiiiiiii" "ppppp" "p" "iiii" "d" "pp" "i" "p" "ii" "p" "ddd" "ii" "d" "i" "d" "pp" "s",
    "rhs_finish": "iiiiiii" "pppppppp" "iii" "ppp" "i" "dddd" "p" "s",
    "rhs_finish_next": "iiiiiii" "pppppppp" "iii" "ppp" "i" "ddddd" "p" "s",
    "reactor_batch": "ii" "pppp" "ppp" "p" "i" "pp" "i" "dddd" "p" "i" "p" "ii" "ii" "d" "pp" "pp" "p" "ii" "s",
}

SIGNATURES: Dict[str, str] = dict(_ARGS)

_CT = {"i": ctypes.c_int, "d": ctypes.c_double, "p": ctypes.c_void_p, "s": ctypes.c_void_p,
       "P": ctypes.c_void_p}


class StreamsError(RuntimeError):
    pass


def as_pointer(x):
    """Address of a torch tensor / numpy array / integer (None -> NULL)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        return x.ctypes.data
    raise TypeError(f"cannot pass {type(x)} as a pointer")


class Library:
    """A shared library exporting the C-ABI under ``prefix`` (``streams_`` for the product)."""

    def __init__(self, path: os.PathLike, prefix: str, mode: int = ctypes.RTLD_GLOBAL):
        self.path = str(path)
        self.prefix = prefix
        self.lib = ctypes.CDLL(self.path, mode=mode)
        self._fn = {}
        for name, kinds in SIGNATURES.items():
            sym = getattr(self.lib, prefix + name, None)
            if sym is None:
                continue
            sym.restype = ctypes.c_int
            sym.argtypes = [_CT[k] for k in kinds]
            self._fn[name] = (sym, kinds)

    def has(self, name: str) -> bool:
        return name in self._fn

    def exported(self) -> List[str]:
        return sorted(self._fn)

    def call(self, name: str, *args):
        sym, kinds = self._fn[name]
        if len(args) != len(kinds):
            raise TypeError(f"{name}: expected {len(kinds)} arguments, got {len(args)}")
        conv = []
        for a, k in zip(args, kinds):
            if k in "psP":
                conv.append(as_pointer(a))
            elif k == "i":
                conv.append(int(a))
            else:
                conv.append(float(a))
        rc = sym(*conv)
        if rc != 0:
            msg = STATUS.get(rc, f"HIP error {rc}")
            raise StreamsError(f"{self.prefix}{name} returned {rc}: {msg}")
        return rc


_LIB = None


def load(path: os.PathLike = LIBPATH) -> Library:
    """Load libstreams_hip.so; raises if the library has not been built."""
    global _LIB
    if _LIB is None:
        if not pathlib.Path(path).exists():
            raise StreamsError(f"{path} is missing: build it with prop-streams2_amd/build.py "
                               "(there is no CPU fallback for the hot path)")
        _LIB = Library(path, "streams_")
        missing = sorted(set(SIGNATURES) - set(_LIB.exported()))
        if missing:
            raise StreamsError(f"libstreams_hip.so lacks entry points: {missing}")
    return _LIB


def kinetics_pack(tables, lib: "Library" = None):
    """Host call of streams_kinetics_pack: the compact per-reaction records (numpy, host) of
    a Tables object's kinetics, to be copied to the device once."""
    import numpy as np
    lib = lib or load()
    sym = lib.lib.streams_kinetics_pack
    sym.restype = ctypes.c_int
    sym.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 11 + [ctypes.c_int]
    t = tables
    f = lambda a, dt=np.float64: np.ascontiguousarray(np.asarray(a, dtype=dt).T)  # noqa: E731
    arrs = [f(t.arr_a), f(t.arr_b), f(t.arr_ea), f(t.tb_eff), f(t.falloff_coeffs), f(t.reac_ty, np.int32),
            f(t.isRev, np.int32), f(t.r_coeffs), f(t.p_coeffs), f(t.mw)]
    need = sym(t.n_species, t.nreactions, *(a.ctypes.data for a in arrs), None, 0)
    if need < 0:
        raise StreamsError(f"streams_kinetics_pack returned {need}: {STATUS.get(need, '')}")
    out = np.zeros(max(need, 1))
    rc = sym(t.n_species, t.nreactions, *(a.ctypes.data for a in arrs), out.ctypes.data, need)
    if rc != need:
        raise StreamsError(f"streams_kinetics_pack returned {rc}: {STATUS.get(rc, '')}")
    return out


def species_supported() -> Tuple[int, ...]:
    lib = load()
    sym = lib.lib.streams_species_supported
    sym.restype = ctypes.c_int
    sym.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    buf = (ctypes.c_int * 64)()
    n = sym(buf, 64)
    return tuple(buf[i] for i in range(min(n, 64)))
