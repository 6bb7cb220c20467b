"""Fortran programs that use the drop-in exactly as a reference caller would (use MatrixSwitch,
call omm) with device storage; they check the reference's own constants and exit non-zero on a
mismatch (example_gamma and example_kpoints: 1e-4, example_sdden: 1e-10)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EX = os.path.join(ROOT, "omm-bundle_amd", "examples")


def run(args, cwd):
    print("CASE", " ".join(args), flush=True)
    p = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120)
    print(p.stdout)
    return p


@pytest.mark.parametrize("op", ["hip", "hmx"])
def test_example_gamma_hip(tmp_path, op):
    p = run([os.path.join(EX, "example_gamma_hip"), os.path.join(ROOT, "tests", "golden", "example_gamma.dat"), op],
            tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "e_min [2]" in p.stdout and "ED_11" in p.stdout


def test_example_sdden_hip(tmp_path):
    p = run([os.path.join(EX, "example_sdden_hip")], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "E_11" in p.stdout


def test_example_kpoints_hip(tmp_path):
    """Two k-points (complex128 H(k), S(k) in HBM), np = 2 per-k state, two SCF steps and ED."""
    g = os.path.join(ROOT, "tests", "golden")
    p = run([os.path.join(EX, "example_kpoints_hip"), os.path.join(g, "example_gamma.dat"),
             os.path.join(g, "example_kpoints.dat"), "gzden"], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "e_min [2,2]" in p.stdout and "ED_11 [2]" in p.stdout


@pytest.mark.parametrize("flavour", ["1", "3"])
def test_example_gamma_hip_flavours(tmp_path, flavour):
    """Cholesky (1) and preconditioned (3) flavours on dense device H and S: same constants."""
    p = run([os.path.join(EX, "example_gamma_hip"), os.path.join(ROOT, "tests", "golden", "example_gamma.dat"),
             "hip", flavour], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "e_min [2]" in p.stdout and "ED_11" in p.stdout
