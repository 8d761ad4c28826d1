"""Host-side mirror of the parts of ICAR's ``options_t`` and ``domain_t`` that
the MPDATA advection path reads.

Arrays are numpy arrays in C order with shape ``(ny, nz, nx)``; that is the
memory layout of ICAR's Fortran column-major ``(ims:ime, kms:kme, jms:jme)``
arrays (i fastest).  ``u``/``jacobian_u`` have ``nx+1`` columns and
``v``/``jacobian_v`` ``ny+1`` rows, as in ``domain_obj.f90:1916-1931``.
"""
from __future__ import annotations

from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np

# src/constants/icar_constants.f90:419-420
kADV_STD = 1
kADV_MPDATA = 2
kFLUXCOR_MONO = 1  # icar_constants.f90:422

# Registration order of advected species, src/objects/domain_obj.f90:68-89.
ADV_VAR_ORDER = ("qv", "theta", "qc", "qr", "qs", "qi", "qg", "ni", "nr", "ns", "ng",
                 "ice1_a", "ice1_c", "ice2_mass", "ice2_number", "ice2_a", "ice2_c",
                 "ice3_mass", "ice3_number", "ice3_a", "ice3_c")


@dataclass
class AdvOptions:
    """``adv_options_type`` (opt_types.f90:119-127) with the namelist defaults
    of ``adv_parameters_namelist`` (options_obj.f90:1579-1584)."""
    boundary_buffer: bool = False  # parsed by the reference but never used
    MPDATA_FCT: bool = True
    mpdata_order: int = 2
    flux_corr: int = 0
    h_order: int = 1
    v_order: int = 1


@dataclass
class Parameters:
    """The ``&parameters`` switches the path reads (options_obj.f90:1025)."""
    advect_density: bool = False
    debug: bool = False


@dataclass
class TimeOptions:
    """``&time_parameters`` (options_obj.f90:2188-2208) with its defaults."""
    RK3: bool = False
    cfl_reduction_factor: float = 0.9
    cfl_strictness: int = 3


@dataclass
class Physics:
    """``&physics adv=`` (options_obj.f90:470-490); 2 selects MPDATA."""
    advection: int = kADV_MPDATA


@dataclass
class Options:
    physics: Physics = field(default_factory=Physics)
    adv_options: AdvOptions = field(default_factory=AdvOptions)
    parameters: Parameters = field(default_factory=Parameters)
    time_options: TimeOptions = field(default_factory=TimeOptions)
    # names requested for advection (options%vars_to_advect); ordered
    vars_to_advect: list = field(default_factory=list)
    # no reference counterpart: 0 exact kernels, 1 fused/batched kernels
    kernel_variant: int = 0

    def advect_vars(self, names) -> None:
        for n in names:
            if n not in self.vars_to_advect:
                self.vars_to_advect.append(n)


def halo_size(options: "Options") -> int:
    """Halo width ICAR gives the grid for the advection options
    (domain_obj.f90:2979-2989 and grid_obj.f90:153-154): ceiling(order/2)
    with order = max(h_order, v_order), plus 2 with the WRF limiter, and 4
    for MPDATA."""
    ao = options.adv_options
    order = max(ao.h_order, ao.v_order)
    if ao.flux_corr == kFLUXCOR_MONO:
        order += 2
    if options.physics.advection == kADV_MPDATA:
        order = 4
    return -(-order // 2)


class Domain:
    """The fields of ``domain_t`` used by ``advect`` and their bounds.

    ``species`` is the ordered ``adv_vars`` registry: name -> (ny,nz,nx)
    float32 array, kept in the order of ``domain_obj.f90:68-89`` by
    :meth:`set_var_lists`.
    """

    FIELDS = ("u", "v", "w", "density", "jacobian", "jacobian_u", "jacobian_v",
              "jacobian_w", "advection_dz")

    def __init__(self, nx: int, ny: int, nz: int, dx: float, ims: int = 1, jms: int = 1,
                 kms: int = 1, halo: int = 2):
        if min(nx, ny, nz) < 3:
            raise ValueError("MPDATA needs at least 3 cells in x, y and z")
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx = float(dx)
        self.ims, self.jms, self.kms = ims, jms, kms
        self.ime, self.jme, self.kme = ims + nx - 1, jms + ny - 1, kms + nz - 1
        # tile bounds as update_with_halos sets them for a one-image domain
        # (grid_obj.f90:238-248); the halo is halo_size(options), 2 for MPDATA
        if halo < 1:
            raise ValueError(f"halo {halo} < 1")
        self.halo = int(halo)
        self.its, self.ite = self.ims + halo, self.ime - halo
        self.jts, self.jte = self.jms + halo, self.jme - halo
        z = np.zeros((ny, nz, nx), np.float32)
        self.u = np.zeros((ny, nz, nx + 1), np.float32)
        self.v = np.zeros((ny + 1, nz, nx), np.float32)
        self.w = z.copy()
        self.density = np.ones_like(z)
        self.jacobian = np.ones_like(z)
        self.jacobian_u = np.ones((ny, nz, nx + 1), np.float32)
        self.jacobian_v = np.ones((ny + 1, nz, nx), np.float32)
        self.jacobian_w = np.ones_like(z)
        self.advection_dz = np.ones_like(z)
        self.species: "OrderedDict[str, np.ndarray]" = OrderedDict()
        self.storage: dict = {}  # name -> (storage enum, pow2 exponent)

    @property
    def shape(self):
        return (self.ny, self.nz, self.nx)

    def field_shape(self, name: str):
        if name in ("u", "jacobian_u"):
            return (self.ny, self.nz, self.nx + 1)
        if name in ("v", "jacobian_v"):
            return (self.ny + 1, self.nz, self.nx)
        return self.shape

    def field_bounds(self, name: str):
        """Fortran (lb, ub) in (i, k, j) order."""
        lb = (self.ims, self.kms, self.jms)
        ub = [self.ime, self.kme, self.jme]
        if name in ("u", "jacobian_u"):
            ub[0] += 1
        if name in ("v", "jacobian_v"):
            ub[2] += 1
        return lb, tuple(ub)

    def set_var_lists(self, options: Options, values: dict) -> None:
        """Builds ``adv_vars`` in the registry order of domain_obj.f90:68-89
        from the requested names; extra names (no ICAR counterpart, e.g. a
        passive tracer) follow in request order."""
        self.species.clear()
        req = list(options.vars_to_advect)
        for name in ADV_VAR_ORDER:
            if name in req:
                self.species[name] = self._as_field(values[name])
        for name in req:
            if name not in self.species:
                self.species[name] = self._as_field(values[name])

    def _as_field(self, a) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.shape != self.shape:
            raise ValueError(f"species shape {a.shape} != {self.shape}")
        return a

    def check(self) -> None:
        for n in self.FIELDS:
            a = getattr(self, n)
            if a.shape != self.field_shape(n) or a.dtype != np.float32 or \
                    not a.flags.c_contiguous:
                raise ValueError(f"field {n}: shape {a.shape} dtype {a.dtype}, expected "
                                 f"{self.field_shape(n)} float32 C-contiguous")
