"""Writes the golden fixtures of the parity tests with the FP64 oracle.

schwarz32.npz: the 32^3 Schwarz-P cell (phi = -(cos X + cos Y + cos Z), Gaussian filter r = 1.5,
beta = 16, volume 0.3, emin = 1e-4), its C from the oracle's PCG at tol 1e-8 and the iteration
count. grey8.npz: an 8^3 grey cell with C at tol 1e-11 and the first matvec / preconditioner
outputs of a seeded random field. The reference's own repository holds no fixtures (its .npz files
are git-ignored), so the oracle writes these; tests/test_oracle.py pins the oracle itself.
Run from the repository root: python tests/golden/make_golden.py
"""
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import linear_ref as O  # noqa: E402

spec = importlib.util.spec_from_file_location(
    "cells", os.path.join(ROOT, "nonlinear-fiber-homogenization_amd", "cells.py"))
cells = importlib.util.module_from_spec(spec)
spec.loader.exec_module(cells)


def main():
    O.set_fft_workers(os.cpu_count() or 1)
    rho = cells.schwarz_cell(32)
    lr = O.LinearRef(rho.shape, emin=1e-4, tol=1e-8)
    C, u, it = lr.C_full(rho)
    np.savez_compressed(os.path.join(HERE, "schwarz32.npz"), rho=rho, C=C, it=it, emin=1e-4, tol=1e-8)
    print("schwarz32", it, C[0, 0], C[0, 1], C[3, 3])

    rng = np.random.default_rng(1234)
    rho8 = cells.exact_symmetrize(cells.density(rng.standard_normal((8, 8, 8)), r=1.0, beta=3.0, v=0.35,
                                                filter_kernel="hat"))
    lr8 = O.LinearRef(rho8.shape, emin=1e-3, tol=1e-11)
    C8, u8, it8 = lr8.C_full(rho8)
    s8 = lr8.stiffness_scale(rho8)
    w = rng.standard_normal((2, 8, 8, 8, 3))
    Kw = np.stack([lr8.apply_K(s8, x) for x in w])
    Mw = lr8.precond(w)
    np.savez_compressed(os.path.join(HERE, "grey8.npz"), rho=rho8, C=C8, it=it8, emin=1e-3, tol=1e-11, w=w, Kw=Kw,
                        Mw=Mw)
    print("grey8", it8)


if __name__ == "__main__":
    main()
