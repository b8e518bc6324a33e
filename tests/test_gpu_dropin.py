"""The Fortran drop-in module (prop-streams2_amd/fortran/streams_kernels_hip.F90) driving the
library from Fortran, as the reference's host code drives its kernel module: the compiled
check program runs on the GPU as a child process and compares the RK bookkeeping and the
periodic halo exchange with host arithmetic."""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
EXE = ROOT / "prop-streams2_amd" / "lib" / "dropin_check"


def _generator():
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen", ROOT / "prop-streams2_amd/fortran/gen_interfaces.py")
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    return gen


def test_generated_module_matches_header():
    """The committed module is, byte for byte, what the generator makes from the current
    header: a stale interface would pass the wrong arguments to an entry point."""
    gen = _generator()
    text = (ROOT / "prop-streams2_amd/fortran/streams_kernels_hip.F90").read_text()
    assert text == gen.render()
    for n in {c for c, _ in gen.parse()}:
        assert f'bind(C, name="{n}")' in text, n


def test_module_interfaces_have_header_arity():
    """Every bind(C) interface declares exactly the header's argument list, in order."""
    import re
    gen = _generator()
    text = (ROOT / "prop-streams2_amd/fortran/streams_kernels_hip.F90").read_text()
    for cname, args in gen.parse():
        m = re.search(r"function " + cname + r"\((.*?)\)\s*&\s*\n\s*bind\(C", text, flags=re.S)
        assert m, cname
        names = [a.strip() for a in m.group(1).replace("&", " ").split(",")]
        assert names == [a for a, _ in args], cname


@pytest.mark.gpu
def test_fortran_dropin_runs_on_gpu():
    assert EXE.exists(), f"{EXE} missing: build with prop-streams2_amd/build.py (needs amdflang)"
    print("[dropin] running", EXE, flush=True)
    res = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "dropin_check: ok" in res.stdout
