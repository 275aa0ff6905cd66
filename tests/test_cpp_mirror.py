"""The C++ solver front end (viennashe::solvers mirror in viennashe-dev_amd/host)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG, ROOT, oracle_system

EXE = os.path.join(ROOT, "tests", "cpp", "test_solver_mirror")


def _build():
    src = os.path.join(ROOT, "tests", "cpp", "test_solver_mirror.cpp")
    if not os.path.exists(EXE) or os.path.getmtime(EXE) < os.path.getmtime(src):
        subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(PKG, "host"), src, "-L", PKG, "-lshesolve",
                        "-Wl,-rpath," + PKG, "-o", EXE], check=True)
    return EXE


def test_cpu_semantics():
    exe = _build()
    r = subprocess.run([exe, "cpu"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "cpu checks passed" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("name,precond", [("diode_L3", "ilu0"), ("mos2d_L3", "bmc_ilu0")])
def test_gpu_solve_through_mirror(tmp_path, name, precond):
    exe = _build()
    s = oracle_system(name)
    A, b, _, _ = s.reduced()
    fin, fout = tmp_path / "sys.bin", tmp_path / "x.bin"
    with open(fin, "wb") as f:
        f.write(np.array([A.n, A.nnz], np.int64).tobytes())
        f.write(A.row_ptr.tobytes())
        f.write(A.col.tobytes())
        f.write(A.val.tobytes())
        f.write(np.ascontiguousarray(b).tobytes())
    r = subprocess.run([exe, "gpu", str(fin), str(fout), precond], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    x = np.fromfile(fout, np.float64)
    res = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
    assert res <= 1e-12
