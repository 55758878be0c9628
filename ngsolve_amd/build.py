"""Build the native parts of ngsolve_amd in-tree.

libngsamd.so    HIP kernels + C ABI (hipcc, gfx950)
_ngsamd*.so     pybind11 host module (g++): FE setup + device handles

Run:  python ngsolve_amd/build.py   (or __graft_entry__.build()); not as
`python -m ngsolve_amd.build`, which would import the unbuilt package first.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NGSAMD_ARCH", "gfx950")

HIP_SOURCES = ["runtime.hip", "spmv.hip", "cg.hip", "matfree.hip", "batch.hip", "blockjacobi.hip", "topology.hip", "multicg.hip", "sparsechol.hip", "mg.hip"]
HOST_SOURCES = ["fe.cpp", "intrule.cpp", "python_module.cpp"]

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_hip(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    lib = os.path.join(HERE, "libngsamd.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, lib)
    return lib


def build_host(verbose=False):
    import pybind11

    os.makedirs(BUILD, exist_ok=True)
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "_ngsamd" + ext)
    tmp = out + ".tmp"
    inc = [
        "-I" + pybind11.get_include(),
        "-I" + sysconfig.get_paths()["include"],
        "-I" + os.path.join(ROOT, "include"),
    ]
    cmd = (["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-fvisibility=hidden"] + inc
           + [os.path.join(HOST, s) for s in HOST_SOURCES]
           + ["-L" + HERE, "-lngsamd", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    o = _run(cmd)
    if verbose and o.strip():
        print(o)
    os.replace(tmp, out)
    return out


def build_adapter_test(verbose=False):
    """The C++ operator layer's test program (tests/cpp), linked against
    libngsamd.so; run by tests/test_adapter.py."""
    src = os.path.join(ROOT, "tests", "cpp", "basematrix_adapter.cpp")
    outdir = os.path.join(ROOT, "tests", "cpp", "build")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "basematrix_adapter")
    tmp = out + ".tmp"
    o = _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), src, "-L" + HERE,
              "-lngsamd", "-Wl,-rpath,$ORIGIN/../../../ngsolve_amd", "-o", tmp])
    if verbose and o.strip():
        print(o)
    os.replace(tmp, out)
    return out


def build(verbose=False):
    lib = build_hip(verbose)
    mod = build_host(verbose)
    build_adapter_test(verbose)
    return lib, mod


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print("built", p)
