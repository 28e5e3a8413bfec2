"""lceqcd_amd -- MI355X (gfx950) lattice update and measurement path of
derek-teaney/lceqcd: 3D SU(2) gauge theory with an adjoint Higgs field.

The HIP library liblceq.so (C ABI in include/lceq.h) owns a device-resident
lattice; :class:`Lattice` mirrors the reference's Fortran routines over it.
"""
from ._lib import LceqError, LIB_PATH
from .batch import Batch
from .ranlux import Ranlux, measinit_key, measinit_seeds
from .mlp import Mlp
from .lattice import (Lattice, SWEEP_KP_OR_METRO, SWEEP_REFERENCE, KC_LINK_KP,
                      KC_LINK_OR, KC_HIGGS, KC_GFIX, couplings_from_xy,
                      couplings_from_trace_convention)

__all__ = ["Lattice", "Batch", "Mlp", "Ranlux", "measinit_key", "measinit_seeds", "LceqError", "LIB_PATH", "SWEEP_KP_OR_METRO", "SWEEP_REFERENCE",
           "KC_LINK_KP", "KC_LINK_OR", "KC_HIGGS", "KC_GFIX", "couplings_from_xy",
           "couplings_from_trace_convention"]
