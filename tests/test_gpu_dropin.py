"""The Fortran drop-in module (prop-streams2_amd/fortran/streams_kernels_hip.F90) driving the
library from Fortran, as the reference's host code drives its kernel module: the compiled
check program runs on the GPU as a child process and compares the RK bookkeeping and the
periodic halo exchange with host arithmetic."""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
EXE = ROOT / "prop-streams2_amd" / "lib" / "dropin_check"


def test_generated_module_matches_header():
    """The committed module is what the generator makes from the current header."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen", ROOT / "prop-streams2_amd/fortran/gen_interfaces.py")
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    names = {c for c, _ in gen.parse()}
    text = (ROOT / "prop-streams2_amd/fortran/streams_kernels_hip.F90").read_text()
    for n in names:
        assert f'bind(C, name="{n}")' in text, n


@pytest.mark.gpu
def test_fortran_dropin_runs_on_gpu():
    assert EXE.exists(), f"{EXE} missing: build with prop-streams2_amd/build.py (needs amdflang)"
    print("[dropin] running", EXE, flush=True)
    res = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "dropin_check: ok" in res.stdout
