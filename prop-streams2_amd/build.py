"""Build recipe for the native pieces of the package.

* ``libstreams_hip.so`` — every ``csrc/*.hip`` compiled for gfx950 with hipcc into one
  shared library with a C-ABI (``include/streams_hip.h``).
* ``libstreams_fortran.so`` + ``dropin_check`` — the Fortran ``iso_c_binding`` interface
  module and its drop-in check program, compiled with amdflang when it is present.

Everything is built in-tree (``prop-streams2_amd/lib``) so that the built files travel
with the source tree.  ``build()`` always recompiles.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
FORTRAN = PKG / "fortran"
ROCM = pathlib.Path(os.environ.get("ROCM_PATH", "/opt/rocm"))

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # keep packed-FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]

# Per-source additions.  finish.hip: machine LICM would hoist the materialised polynomial
# constants of exp/log out of the chemistry's reaction loop into ~35 VGPRs that stay live
# through it; without the hoist the species kernel fits 6 waves per SIMD instead of 4.
FILE_FLAGS = {
    "finish.hip": ["-mllvm", "-disable-machine-licm"],
    "euler.hip": ["-ffp-contract=off"],
}


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(map(str, cmd))}")
    return res


def hipcc() -> str:
    p = ROCM / "bin" / "hipcc"
    return str(p) if p.exists() else "hipcc"


def build_hip(jobs: int = 8, verbose: bool = False) -> pathlib.Path:
    LIB.mkdir(exist_ok=True)
    objdir = LIB / "obj"
    objdir.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError("no HIP sources found")

    def compile_one(src: pathlib.Path) -> pathlib.Path:
        obj = objdir / (src.stem + ".o")
        cmd = [hipcc(), *HIPCC_FLAGS, *FILE_FLAGS.get(src.name, []), "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        return obj

    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(sources)))) as ex:
        objs = list(ex.map(compile_one, sources))
    out = LIB / "libstreams_hip.so"
    tmp = LIB / "libstreams_hip.so.tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    return out


def build_fortran(verbose: bool = False):
    """Compile the Fortran drop-in module and check program (amdflang)."""
    fc = shutil.which("amdflang") or (str(ROCM / "bin" / "amdflang") if (ROCM / "bin" / "amdflang").exists() else None)
    if fc is None:
        return None
    LIB.mkdir(exist_ok=True)
    mod_src = FORTRAN / "streams_kernels_hip.F90"
    prog_src = FORTRAN / "dropin_check.F90"
    if not mod_src.exists():
        return None
    moddir = LIB / "fmod"
    moddir.mkdir(exist_ok=True)
    obj = LIB / "streams_kernels_hip.o"
    _run([fc, "-O2", "-fPIC", "-cpp", f"-I{FORTRAN}", "-module-dir", str(moddir), "-c", str(mod_src), "-o", str(obj)])
    outs = [obj]
    if prog_src.exists():
        exe = LIB / "dropin_check"
        _run([fc, "-O2", "-cpp", f"-I{FORTRAN}", f"-I{moddir}", str(prog_src), str(obj), "-o", str(exe),
              f"-L{LIB}", "-lstreams_hip", f"-Wl,-rpath,{LIB}", f"-Wl,-rpath,{ROCM / 'lib'}",
              f"-L{ROCM / 'lib'}", "-lamdhip64"])
        outs.append(exe)
    return outs


def build(verbose: bool = False):
    so = build_hip(verbose=verbose)
    build_fortran(verbose=verbose)
    return so


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
