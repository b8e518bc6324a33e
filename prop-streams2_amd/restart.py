"""Restart files of the reference for one rank (field.F90:349-477, multideal.F90:2381-2487).

* ``rst.bin`` (io_type 2, written through an MPI-IO subarray view, field.F90:427-477): no
  record markers; variable after variable, each the global (nx, ny, nz) interior in Fortran
  order.  On one rank that is the dense interior block w(1:nx,1:ny,1:nz,1:nv), which the
  device extracts from the padded field (``streams_field_interior``).
* ``rst0_/rst1_XXXX_XXXX_XXXX.bin`` (io_type 1, field.F90:349-361, 412-425): one Fortran
  sequential unformatted record of the host layout w(1:nv,1:nx,1:ny,1:nz) (variable
  fastest), framed by 4-byte length markers; rst0 is read, rst1 written, XXXX the rank's
  Cartesian coordinates.  The device transposes the field to that layout
  (``streams_copy_gpu_cpu``).  Records above 2**31 - 1 bytes would need subrecords and are
  refused: such fields use rst.bin.
* ``field_info0.dat`` / ``field_info1.dat`` (write_field_info / read_field_info): the
  cfgio format of the reference's third-party parser, section ``[field_info]`` with
  ``key = value`` lines (the cycle, time and output counters of the run).
"""
from __future__ import annotations

import pathlib
import struct
from typing import Dict

import numpy as np

FIELD_INFO_KEYS = ("icyc0", "time0", "itav", "itslice_vtr", "time_from_last_rst", "time_from_last_write",
                   "time_from_last_stat", "time_from_last_slice", "time_from_last_slice_vtr", "istore")
INT_KEYS = {"icyc0", "itav", "itslice_vtr", "istore"}
MAX_RECORD = 2 ** 31 - 1


def serial_name(prefix: str, ncoords=(0, 0, 0)) -> str:
    """rst0_XXXX_XXXX_XXXX.bin / rst1_... (the I4.4 rank coordinates of field.F90:353-358)."""
    return f"{prefix}_{ncoords[0]:04d}_{ncoords[1]:04d}_{ncoords[2]:04d}.bin"


def write_rst_bin(path, interior: np.ndarray):
    """interior: (nv, nz, ny, nx) C-ordered, i.e. w(1:nx,1:ny,1:nz,1:nv) in Fortran order."""
    np.ascontiguousarray(interior, dtype="<f8").tofile(str(path))


def read_rst_bin(path, nx: int, ny: int, nz: int, nv: int) -> np.ndarray:
    a = np.fromfile(str(path), dtype="<f8")
    if a.size != nx * ny * nz * nv:
        raise ValueError(f"{path}: {a.size} values, expected {nx * ny * nz * nv} for ({nx},{ny},{nz},{nv})")
    return a.reshape(nv, nz, ny, nx)


def write_rst_serial(path, host_interior: np.ndarray):
    """host_interior: (nz, ny, nx, nv) C-ordered, i.e. w(1:nv,1:nx,1:ny,1:nz) in Fortran order."""
    data = np.ascontiguousarray(host_interior, dtype="<f8").tobytes()
    if len(data) > MAX_RECORD:
        raise ValueError(f"record of {len(data)} bytes exceeds one 4-byte-marked record: write rst.bin instead")
    with open(path, "wb") as fh:
        fh.write(struct.pack("<i", len(data)))
        fh.write(data)
        fh.write(struct.pack("<i", len(data)))


def read_rst_serial(path, nx: int, ny: int, nz: int, nv: int) -> np.ndarray:
    raw = pathlib.Path(path).read_bytes()
    (n0,) = struct.unpack_from("<i", raw, 0)
    (n1,) = struct.unpack_from("<i", raw, len(raw) - 4)
    if n0 != n1 or n0 != len(raw) - 8:
        raise ValueError(f"{path}: not one sequential unformatted record (markers {n0}, {n1}, size {len(raw)})")
    if n0 != 8 * nx * ny * nz * nv:
        raise ValueError(f"{path}: record of {n0} bytes, expected {8 * nx * ny * nz * nv}")
    return np.frombuffer(raw, dtype="<f8", count=n0 // 8, offset=4).reshape(nz, ny, nx, nv).copy()


def _fortran_value(v) -> str:
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    # list-directed style with 17 significant digits: read back bit for bit
    m, e = f"{float(v):.16E}".split("E")
    return f"{m}E{int(e):+04d}"


def write_field_info(path, info: Dict[str, float]):
    lines = ["[field_info]"]
    for k in FIELD_INFO_KEYS:
        if k in info:
            lines.append(f"{k} = {_fortran_value(info[k])}")
    for k, v in info.items():
        if k not in FIELD_INFO_KEYS:
            lines.append(f"{k} = {_fortran_value(v)}")
    pathlib.Path(path).write_text("\n".join(lines) + "\n")


def read_field_info(path) -> Dict[str, float]:
    out: Dict[str, float] = {}
    for raw in pathlib.Path(path).read_text().splitlines():
        line = raw.split(";", 1)[0].split("#", 1)[0].strip()
        if not line or line.startswith("["):
            continue
        k, v = (x.strip() for x in line.split("=", 1))
        v = v.replace("d", "e").replace("D", "E")
        out[k] = int(float(v)) if k in INT_KEYS else float(v)
    if "time_from_last_slice_vtr" not in out and "time_from_last_slice" in out:
        out["time_from_last_slice_vtr"] = out["time_from_last_slice"]       # read_field_info's default
    return out
