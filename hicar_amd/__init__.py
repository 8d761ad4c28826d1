"""MI355X-native MPDATA advection for ICAR (adv=2).

The product is the HIP library ``libicar_adv.so`` (C-ABI in
``include/icar_adv.h``); this package holds its build script, a ctypes
binding, and a host-side mirror of ICAR's advection interface
(``adv_var_request``, ``adv_init``, ``advect``).
"""
from .domain import Domain, Options, kADV_MPDATA, kADV_STD  # noqa: F401

__all__ = ["Domain", "Options", "kADV_MPDATA", "kADV_STD", "advection", "synthetic"]
