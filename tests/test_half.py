"""Host rounding of the reference gradients to the 16-bit MFMA operand
formats (ngsolve_amd/csrc/half.hpp) against numpy's IEEE conversions."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r"""
#include <cstdio>
#include <cstdint>
#include "half.hpp"
int main(int argc, char** argv) {
  FILE* f = std::fopen(argv[1], "rb");
  FILE* o = std::fopen(argv[2], "wb");
  float v;
  while (std::fread(&v, 4, 1, f) == 1) {
    uint16_t h[2] = {ngsamd::ToF16(v), ngsamd::ToBf16(v)};
    std::fwrite(h, 2, 2, o);
  }
  std::fclose(f);
  std::fclose(o);
  return 0;
}
"""


@pytest.fixture(scope="module")
def conv(tmp_path_factory):
    d = tmp_path_factory.mktemp("half")
    src, exe = d / "t.cpp", d / "t"
    src.write_text(PROG)
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "ngsolve_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)

    def run(x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        (d / "in").write_bytes(x.tobytes())
        subprocess.run([str(exe), str(d / "in"), str(d / "out")], check=True)
        return np.frombuffer((d / "out").read_bytes(), dtype=np.uint16).reshape(-1, 2)
    return run


def test_f16_matches_numpy(conv):
    rng = np.random.default_rng(0)
    # normals, subnormals, halfway cases, overflow, zeros, inf, nan
    x = np.concatenate([
        rng.standard_normal(20000) * 10.0 ** rng.integers(-9, 6, 20000),
        np.array([0.0, -0.0, 65504.0, 65519.0, 65520.0, 1e6, -1e6, np.inf, -np.inf, 6e-8, 2.9e-8, 3e-8, 5.96e-8]),
        # exact halfway points between neighbouring halves
        ((np.arange(1, 2048, dtype=np.float32) + 0.5) * np.float32(2.0 ** -10)),
        ((np.arange(1, 1024, dtype=np.float32) + 0.5) * np.float32(2.0 ** -24)),
    ]).astype(np.float32)
    got = conv(x)[:, 0]
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, ref)
    nan = conv(np.array([np.nan], dtype=np.float32))[0, 0]
    assert np.isnan(np.array([nan], dtype=np.uint16).view(np.float16)[0])


def test_bf16_round_to_nearest_even(conv):
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000)).astype(np.float32)
    got = conv(x)[:, 1].astype(np.uint32) << 16
    back = got.view(np.float32)
    # nearest: the error is at most half a bf16 ulp, and ties go to even
    u = x.view(np.uint32)
    lo = (u & 0xFFFF0000).view(np.float32)
    hi = ((u & 0xFFFF0000) + 0x10000).view(np.float32)
    dlo, dhi = np.abs(x - lo), np.abs(hi - x)
    expect = np.where(dlo < dhi, lo, np.where(dhi < dlo, hi, np.where((u >> 16) & 1, hi, lo)))
    assert np.array_equal(back, expect)


SPLIT_PROG = r"""
#include <cstdio>
#include <cstdint>
#include "half.hpp"
int main(int argc, char** argv) {
  FILE* f = std::fopen(argv[1], "rb");
  FILE* o = std::fopen(argv[2], "wb");
  double v;
  while (std::fread(&v, 8, 1, f) == 1) {
    uint16_t hl[2];
    ngsamd::SplitF16(v, &hl[0], &hl[1]);
    const double back[2] = {ngsamd::F16ToDouble(hl[0]), ngsamd::F16ToDouble(hl[1])};
    std::fwrite(back, 8, 2, o);
  }
  std::fclose(f);
  std::fclose(o);
  return 0;
}
"""


@pytest.fixture(scope="module")
def split(tmp_path_factory):
    d = tmp_path_factory.mktemp("split")
    src, exe = d / "s.cpp", d / "s"
    src.write_text(SPLIT_PROG)
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "ngsolve_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)

    def run(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        (d / "in").write_bytes(x.tobytes())
        subprocess.run([str(exe), str(d / "in"), str(d / "out")], check=True)
        return np.frombuffer((d / "out").read_bytes(), dtype=np.float64).reshape(-1, 2)
    return run


def test_f16_split(split):
    """Split FP16 operands: hi = fp16(v) decodes to numpy's float16 value, lo
    = fp16(v - hi), and hi + lo carries about 22 significant bits (absolute
    2^-25 below the normal range, where lo is subnormal)."""
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-6, 4, 20000),
                        rng.uniform(-1, 1, 5000), [0.0, 1.0, -1.0, 65504.0, 2.0 ** -14, 2.0 ** -24]])
    hl = split(x)
    hi_ref = x.astype(np.float32).astype(np.float16).astype(np.float64)
    assert np.array_equal(hl[:, 0], hi_ref)
    lo_ref = (x - hi_ref).astype(np.float32).astype(np.float16).astype(np.float64)
    assert np.array_equal(hl[:, 1], lo_ref)
    err = np.abs(x - hl[:, 0] - hl[:, 1])
    assert np.all(err <= 2.0 ** -22 * np.abs(x) + 2.0 ** -25)
    # every finite FP16 value decodes exactly (round trip through the split)
    allh = np.arange(0, 65536, dtype=np.uint32).astype(np.uint16).view(np.float16).astype(np.float64)
    fin = np.isfinite(allh)
    hl = split(allh[fin])
    assert np.array_equal(hl[:, 0], allh[fin]) and np.all(hl[:, 1] == 0)
