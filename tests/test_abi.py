"""The C-ABI library loads and exports every symbol include/ngsamd.h declares.
CPU only: no compute call is made."""
import ctypes
import os
import re

from conftest import ROOT


def _declared():
    with open(os.path.join(ROOT, "include", "ngsamd.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(ngsamd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ["ngsamd_init", "ngsamd_csr_create", "ngsamd_op_mult_add", "ngsamd_mf_create", "ngsamd_cg_solve",
                 "ngsamd_ir_cg_solve", "ngsamd_jacobi_create", "ngsamd_vec_dot", "ngsamd_mf_element_grad"]:
        assert must in names


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(os.path.join(ROOT, "ngsolve_amd", "libngsamd.so"))
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_last_error_without_init():
    lib = ctypes.CDLL(os.path.join(ROOT, "ngsolve_amd", "libngsamd.so"))
    lib.ngsamd_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.ngsamd_last_error(), bytes)
    # an entry point used before ngsamd_init reports a state error, no crash
    lib.ngsamd_synchronize.restype = ctypes.c_int
    assert lib.ngsamd_synchronize() == 3


def test_python_package_imports_native_modules():
    import ngsolve_amd
    assert ngsolve_amd._N.__file__.endswith(".so")


def test_precision_names_match_header():
    """The Python precision names map to the header's NGSAMD_PREC_* values."""
    with open(os.path.join(ROOT, "include", "ngsamd.h")) as f:
        consts = {m.group(1): int(m.group(2)) for m in re.finditer(r"#define NGSAMD_PREC_(\w+) (\d+)", f.read())}
    from ngsolve_amd._native import PREC
    assert consts == {"F64": 0, "F32": 1, "BF16": 2, "F16": 3, "F16X3": 4}
    for name, key in [("f64", "F64"), ("f32", "F32"), ("bf16", "BF16"), ("f16", "F16"), ("f16x3", "F16X3")]:
        assert PREC[name] == consts[key]
