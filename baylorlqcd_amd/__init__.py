"""MI355X-native Wilson / twisted-mass solve path of BaylorLQCD "qqcd".

HIP library ``libqqcdhip.so`` (C ABI in ``include/qqcd_hip.h``) plus Python
mirrors of the reference's module procedures:

* ``diracops``  -- Hsingle, Hdouble, Hdbletm, gammamult, gam5x
* ``inverters`` -- gmresdr, psc, gmresproject, bicgstab, vecdot
* ``quark``     -- pointsource, fermprop
* ``lattice``   -- library start-up and the single-process geometry tables
* ``gauge``     -- seeded SU(3) fields
"""
from . import _lib, gauge, lattice  # noqa: F401

__all__ = ["_lib", "gauge", "lattice"]
