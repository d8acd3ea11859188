"""The C-ABI library loads (no GPU needed) and exports every symbol that include/nlfh.h declares."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "nlfh.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nlfh_[A-Za-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.load_library()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_binding_covers_header(pkg):
    from importlib import import_module  # noqa: F401
    bound = {name for name, _, _ in pkg._lib.SIGNATURES}
    assert bound == set(_declared())


def test_error_path_without_device(pkg):
    lib = pkg.load_library()
    h = ctypes.c_void_p()
    # invalid arguments are rejected before any device call
    assert lib.nlfh_plan_create(None, 1, 1.0, 0.3, 6, ctypes.byref(h)) != 0
    assert b"ncells" in lib.nlfh_last_error()
    shapes = (ctypes.c_int * 3)(7, 8, 8)  # 7 is not a product of 2, 3, 5
    assert lib.nlfh_version() == 1
    assert lib.nlfh_plan_bytes(None) == 0
    del shapes


def test_jacobi_bound_is_the_element_bound(pkg):
    """nlfh_jacobi_bound (host arithmetic only) is lambda_max(Ke) / Ke_00 of the oracle's Ke for several Poisson
    ratios, and rejects invalid materials."""
    from oracle import linear_ref as O
    lib = pkg.load_library()
    for nu in (0.0, 0.2, 0.3, 0.45):
        assert abs(lib.nlfh_jacobi_bound(2.0, nu) - O.jacobi_bound(2.0, nu)) < 1e-11 * O.jacobi_bound(2.0, nu)
    assert lib.nlfh_jacobi_bound(1.0, 0.5) == -1.0 and lib.nlfh_jacobi_bound(-1.0, 0.3) == -1.0
