"""qem4molly_amd -- MI355X (gfx950) implementation of QEM4Molly's quasi-Ewald
force path behind the reference's interaction API.

    from qem4molly_amd import QEM_inter, System, Atom, forces

The compute runs in libqem_hip.so (C ABI, include/qem.h); this package is the
host-side mirror of the reference's Molly interface plus a ctypes binding.
"""
from ._lib import (Context, DeviceArray, QEMError, gauss_legendre, inter_params,
                   synchronize, PREC_FP64, PREC_MIXED, SEM_SWAP_LIMITS, SEM_MINIMAGE_DIR,
                   SEM_NEWTON_IMAGES, SEM_CORRECTED, SEM_POLE_SPLIT, LONG_NONSORT, LONG_SORT, LONG_SELF,
                   LONG_K0, LONG_ALL, Wall, wall, WALL_SEM_ABS_CUTOFF, WALL_SEM_TRUE_FORCE)
from .interactions import (Atom, System, QEM_short, QEM_long, QEM_inter, forces, F_short_ij,
                           F_short_i, K_set_generator, greens_element, greens_element_init,
                           Fl_non_sort, Fl_sort, Flz_sort_k0, Flz_self, F_l_total, F_long,
                           Substrate_LJ)

__all__ = [
    "Context", "DeviceArray", "QEMError", "gauss_legendre", "inter_params", "synchronize",
    "PREC_FP64", "PREC_MIXED", "SEM_SWAP_LIMITS", "SEM_MINIMAGE_DIR", "SEM_NEWTON_IMAGES",
    "SEM_CORRECTED", "SEM_POLE_SPLIT", "LONG_NONSORT", "LONG_SORT", "LONG_SELF", "LONG_K0", "LONG_ALL",
    "Atom", "System", "QEM_short", "QEM_long", "QEM_inter", "forces", "F_short_ij",
    "F_short_i", "K_set_generator", "greens_element", "greens_element_init", "Fl_non_sort",
    "Fl_sort", "Flz_sort_k0", "Flz_self", "F_l_total", "F_long", "Substrate_LJ", "Wall",
    "wall", "WALL_SEM_ABS_CUTOFF", "WALL_SEM_TRUE_FORCE",
]
