"""The C++ operator layer (include/ngsamd_la.hpp: BaseMatrix's virtual
interface over the C ABI) through its test program tests/cpp/
basematrix_adapter.cpp, built by the package build: on the CPU the program
links and maps a missing ngsamd_init to an exception; on the GPU it checks
DevSparseMatrix Mult / MultAdd, the BaseMatrix defaults and their recursion
guard, SumMatrix / ScaleMatrix, CGSolver with Jacobi and block Jacobi against
a host PCG in the program, and error reporting."""
import os
import subprocess

import pytest

from conftest import ROOT

EXE = os.path.join(ROOT, "tests", "cpp", "build", "basematrix_adapter")


def _run(*args):
    assert os.path.exists(EXE), "build the package first (tests/cpp/build/basematrix_adapter is missing)"
    return subprocess.run([EXE, *args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)


def test_adapter_links_and_reports_state_error():
    r = _run("--no-device")
    assert r.returncode == 0, r.stdout
    assert "basematrix adapter (no device): OK" in r.stdout


@pytest.mark.gpu
def test_adapter_on_device():
    r = _run()
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout
    assert "basematrix adapter: OK" in r.stdout
