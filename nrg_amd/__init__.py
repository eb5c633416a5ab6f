"""MI355X (gfx950) chemistry step for NRG: Gear-BDF stiff kinetics per cell in HIP.

The product is ``libnrg_chem.so`` (C-ABI, include/nrg_chem.h); this package is
its Python host side: the mechanism reader, a ctypes view of the ABI and a
mirror of NRG's ``chemical_kinetics_solver`` interface.
"""
from .mechanism import Mechanism, load_mechanism  # noqa: F401

__all__ = ["Mechanism", "load_mechanism"]
