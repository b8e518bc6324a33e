"""CPU side of the parity tests: the C restatement in oracle/ behind the same call table
as the HIP library (entry points prefixed orc_ instead of streams_), on numpy arrays.

Only tests, smoke() and bench.py's cpu_baseline leg use this module.
"""
import ctypes
import pathlib
import subprocess

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"


def build_oracle():
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True, capture_output=True)


def load_oracle(variant: str = "liboracle.so"):
    import prop_streams2_amd.capi as capi
    path = ORACLE_DIR / "lib" / variant
    if not path.exists():
        build_oracle()
    return capi.Library(path, "orc_", mode=ctypes.RTLD_LOCAL)


class OracleBackend:
    """numpy arrays + liboracle.so, with the backend interface the Solver uses."""

    name = "oracle"
    stream = None

    def __init__(self, lib=None):
        self.lib = lib or load_oracle()

    def zeros(self, shape, dtype="f8"):
        return np.zeros(shape, dtype={"f8": np.float64, "i4": np.int32}[dtype])

    def asarray(self, a, dtype="f8"):
        return np.ascontiguousarray(a, dtype={"f8": np.float64, "i4": np.int32}[dtype]).copy()

    def to_numpy(self, x):
        return np.array(x, copy=True)

    def copy_(self, dst, src):
        np.copyto(dst, src)

    def sync(self):
        pass

    def call(self, name, *args):
        return self.lib.call(name, *args)

    def has(self, name):
        return self.lib.has(name)
