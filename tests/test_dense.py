"""Host dense inverse of the block-Jacobi setup (ngsolve_amd/csrc/dense.hpp)
against numpy.linalg.inv: random, pivoting-dependent (zero leading
diagonal), 1x1 and singular matrices."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r"""
#include <cstdio>
#include <vector>
#include "dense.hpp"
int main(int argc, char** argv) {
  FILE* f = std::fopen(argv[1], "rb");
  FILE* o = std::fopen(argv[2], "wb");
  int m;
  while (std::fread(&m, 4, 1, f) == 1) {
    std::vector<double> a((size_t)m * m);
    if (std::fread(a.data(), 8, a.size(), f) != a.size()) return 1;
    const int ok = ngsamd::InvertInPlace(a.data(), m) ? 1 : 0;
    std::fwrite(&ok, 4, 1, o);
    std::fwrite(a.data(), 8, a.size(), o);
  }
  std::fclose(f);
  std::fclose(o);
  return 0;
}
"""


@pytest.fixture(scope="module")
def inverter(tmp_path_factory):
    d = tmp_path_factory.mktemp("dense")
    src, exe = d / "t.cpp", d / "t"
    src.write_text(PROG)
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "ngsolve_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)

    def run(mats):
        buf = b"".join(np.int32(len(m)).tobytes() + np.ascontiguousarray(m, dtype=np.float64).tobytes()
                       for m in mats)
        (d / "in").write_bytes(buf)
        subprocess.run([str(exe), str(d / "in"), str(d / "out")], check=True)
        out, pos, res = (d / "out").read_bytes(), 0, []
        for m in mats:
            k = len(m)
            ok = int(np.frombuffer(out, dtype=np.int32, count=1, offset=pos)[0])
            inv = np.frombuffer(out, dtype=np.float64, count=k * k, offset=pos + 4).reshape(k, k)
            res.append((ok, inv))
            pos += 4 + 8 * k * k
        return res
    return run


def test_inverse_matches_numpy(inverter):
    rng = np.random.default_rng(0)
    mats = [rng.standard_normal((m, m)) + (m if s else 0) * np.eye(m) for m in (1, 2, 3, 5, 8, 17, 40)
            for s in (0, 1)]
    # pivoting needed: zero leading diagonal, permutation-like matrices
    mats.append(np.array([[0.0, 2.0, 1.0], [1.0, 0.0, 3.0], [4.0, 1.0, 0.0]]))
    mats.append(np.eye(6)[rng.permutation(6)] * np.arange(1, 7))
    for (ok, inv), a in zip(inverter(mats), mats):
        assert ok == 1
        ref = np.linalg.inv(a)
        assert np.abs(inv - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max())
        assert np.abs(a @ inv - np.eye(len(a))).max() < 1e-10


def test_singular_is_reported(inverter):
    a = np.array([[1.0, 2.0], [2.0, 4.0]])
    z = np.zeros((3, 3))
    (ok1, _), (ok2, _) = inverter([a, z])
    assert ok2 == 0
    # exactly singular 2x2: the second pivot is exactly 0
    assert ok1 == 0
