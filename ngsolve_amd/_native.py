"""Loads the native modules.  There is no fallback: importing ngsolve_amd
without the built HIP library fails with an ImportError naming the build
step."""
import os

# torch, when installed, ships its own HIP runtime under the soname of the
# system one: loaded first, it is the one libngsamd binds to, so the process
# has a single runtime and torch tensors can be handed to the library
# (ngsolve_amd.autograd); loaded after libngsamd, torch finds no device
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    pass

_HERE = os.path.dirname(os.path.abspath(__file__))

try:
    from . import _ngsamd as N  # noqa: F401  (links libngsamd.so through its rpath)
except ImportError as e:  # pragma: no cover - exercised only when unbuilt
    raise ImportError(
        "ngsolve_amd native modules are not built (libngsamd.so / _ngsamd*.so); "
        "run `python ngsolve_amd/build.py` first"
    ) from e

LIBRARY_PATH = os.path.join(_HERE, "libngsamd.so")

_initialised = False


def ensure_device(device=0):
    """Initialise libngsamd on `device` once per process."""
    global _initialised
    if not _initialised:
        N.init(device)
        _initialised = True
    return N


PREC = {"f64": 0, "fp64": 0, "float64": 0, "f32": 1, "fp32": 1, "float32": 1, "bf16": 2, "bfloat16": 2, "f16": 3, "fp16": 3, "float16": 3, "f16x3": 4}
SCATTER = {"gather": 0, "atomic": 1}
