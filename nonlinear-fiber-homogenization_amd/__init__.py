"""MI355X-native small-strain periodic homogenization (drop-in for nlfiber/linear.py).

The directory name is not a Python identifier; load it under the module name `nlfh_amd`
(see `__graft_entry__.load_package`). Public API:
  LinearHomogenization, Plan            (linear.py: the reference's class on libnlfh.so)
  BatchedHomogenization                 (batch.py: many cells x 6 load cases per launch)
  cells                                 (synthetic cubic-symmetric unit cells)
  mech                                  (finite-strain Cell / Mixed Newton-CG: nlfiber/mech.py)
  cubic_invariants, cubic_leakage, hashin_shtrikman, iso_D, lame, unit_strain, ...
"""
from . import cells  # noqa: F401
from ._lib import NlfhError, load as load_library  # noqa: F401
from .linear import (  # noqa: F401
    GPTS, NODES, VOIGT, B_matrices, LinearHomogenization, Plan, clip01, cubic_invariants, cubic_leakage, dN_ref,
    from_planar, hashin_shtrikman, iso_D, lame, to_planar, unit_strain,
)

__version__ = "0.1.0"
from .batch import BatchedHomogenization  # noqa: E402,F401
from . import learned  # noqa: E402,F401
from . import density, fiber, mech, topology  # noqa: E402,F401
