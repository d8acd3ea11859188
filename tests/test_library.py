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


def test_sanitized_host_abi_checks(tmp_path):
    """The C ABI built with AddressSanitizer and UndefinedBehaviorSanitizer on its host code
    (tests/sanitize/build.sh) runs the host check program tests/sanitize/abi_errors.c clean: argument refusals
    of every entry-point family and the host arithmetic (the Jacobi spectral bound, workspace sizes), no device
    calls, no sanitizer report, leak checking on."""
    import shutil
    import subprocess
    import pytest
    if shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc is not available")
    b = subprocess.run(["bash", os.path.join(ROOT, "tests", "sanitize", "build.sh"), str(tmp_path)], capture_output=True,
                       text=True, timeout=900)
    assert b.returncode == 0, b.stderr[-2000:]
    r = subprocess.run([str(tmp_path / "abi_errors")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "all checks passed" in r.stdout
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
