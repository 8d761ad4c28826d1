"""Host-side mirror of ICAR's advection module (MPDATA and the standard scheme).

``adv_var_request``, ``adv_init`` and ``advect`` keep the names, arguments and
behaviour of ``src/physics/advection_driver.f90:23-126``; the work happens on
the GPU through the C-ABI of ``libicar_adv.so``.  Species, winds and metric
stay resident on the device between calls; ``advect`` updates the host
arrays in place (as the reference does) unless ``sync_host=False``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .domain import Domain, Options, kADV_MPDATA, kADV_STD

_fp = ctypes.POINTER(ctypes.c_float)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)


class Advection:
    """One device-resident advection state for one domain (one C-ABI
    handle): MPDATA (adv=2) or the standard scheme (adv=1), as the options
    select."""

    def __init__(self, domain: Domain, options: Options):
        domain.check()
        self.lib = _lib.load()
        self.domain = domain
        self.options = options
        g = _lib.Grid(domain.ims, domain.ime, domain.kms, domain.kme, domain.jms, domain.jme,
                      domain.its, domain.ite, domain.jts, domain.jte, domain.dx)
        ao = options.adv_options
        o = _lib.Opts(int(ao.mpdata_order), int(ao.MPDATA_FCT),
                      int(options.parameters.advect_density), int(options.parameters.debug),
                      int(options.time_options.RK3), int(options.kernel_variant),
                      int(options.physics.advection), int(ao.h_order), int(ao.v_order),
                      int(ao.flux_corr))
        h = ctypes.c_void_p()
        _lib.check(self.lib.icar_adv_create(ctypes.byref(g), ctypes.byref(o), ctypes.byref(h)))
        self.h = h
        self._keep = []
        for fid, name in enumerate(_lib.FIELD_NAMES):
            arr = getattr(domain, name)
            lb, ub = domain.field_bounds(name)
            lba = (ctypes.c_int * 3)(*lb)
            uba = (ctypes.c_int * 3)(*ub)
            self._check(self.lib.icar_adv_bind_field(self.h, fid, _ptr(arr), lba, uba))
            self._keep.append(arr)
        self.names = list(domain.species.keys())
        if len(self.names) > _lib.MAX_SPECIES:
            raise ValueError(f"at most {_lib.MAX_SPECIES} advected species")
        for idx, name in enumerate(self.names):
            arr = domain.species[name]
            storage, e = domain.storage.get(name, (_lib.STORE_F32, 0))
            self._check(self.lib.icar_adv_bind_species(self.h, idx, _ptr(arr), storage, e))
            self._keep.append(arr)
        self.upload()

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc: int) -> None:
        _lib.check(rc, self.h)

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.icar_adv_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- residency --------------------------------------------------------
    def upload(self, fields: bool = True, species: bool = True) -> None:
        mask = (_lib.MASK_ALL_FIELDS if fields else 0) | (_lib.MASK_ALL_SPECIES if species else 0)
        self._check(self.lib.icar_adv_upload(self.h, mask))

    def download(self, species: bool = True, fields: bool = False) -> None:
        mask = (_lib.MASK_ALL_FIELDS if fields else 0) | (_lib.MASK_ALL_SPECIES if species else 0)
        self._check(self.lib.icar_adv_download(self.h, mask))

    # -- the step ----------------------------------------------------------
    def advect(self, dt: float) -> None:
        self._check(self.lib.icar_adv_advect(self.h, ctypes.c_float(dt)))

    def apply_bc(self, bc_x: int, bc_y: int) -> None:
        self._check(self.lib.icar_adv_apply_bc(self.h, bc_x, bc_y))

    def run(self, dt: float, nsteps: int, bc_x: int = 0, bc_y: int = 0) -> None:
        self._check(self.lib.icar_adv_run(self.h, ctypes.c_float(dt), nsteps, bc_x, bc_y))

    def sync(self) -> None:
        self._check(self.lib.icar_adv_sync(self.h))

    # -- diagnostics ---------------------------------------------------------
    def mass_budget(self) -> np.ndarray:
        out = np.zeros(len(self.names), np.float64)
        self._check(self.lib.icar_adv_mass_budget(
            self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def max_courant(self) -> np.ndarray:
        out = np.zeros(3, np.float64)
        self._check(self.lib.icar_adv_max_courant(
            self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def set_compensation(self, on: bool) -> None:
        """The fused kernels' flux divergence with compensated sums (not the
        reference's operation order) or, by default, without."""
        self._check(self.lib.icar_adv_set_compensation(self.h, int(bool(on))))

    def divergence(self) -> tuple[int, float]:
        """test_divergence (adv_mpdata.f90:445-475) on the Courant numbers of
        the last step: the number of interior cells with |div| > 1e-3 and the
        largest |div|."""
        out = np.zeros(2, np.float64)
        self._check(self.lib.icar_adv_divergence(
            self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return int(out[0]), float(out[1])

    def balance_w(self) -> None:
        """balance_uvw (wind.f90:88-164) on the device copies of the winds:
        the device w becomes the continuity-balanced one; ``download(
        fields=True)`` returns it to the host."""
        self._check(self.lib.icar_adv_balance_w(self.h))

    def compute_dt(self, cfl: float, cfl_strictness: int, dz_levels) -> float:
        """compute_dt (time_step.f90:51-168, use_density=.false.) on the device
        winds; raises IcarAdvError (code ECOURANT) when dt < 0.1 s, where the
        reference stops."""
        dzl = np.ascontiguousarray(dz_levels, dtype=np.float32)
        if dzl.shape != (self.domain.nz,):
            raise ValueError(f"dz_levels needs {self.domain.nz} values")
        dt = ctypes.c_float()
        self._check(self.lib.icar_adv_compute_dt(self.h, ctypes.c_float(cfl), int(cfl_strictness),
                                                 _ptr(dzl), ctypes.byref(dt)))
        return dt.value

    def set_relax_filter(self, west: bool, east: bool, north: bool, south: bool,
                         halo: int = 2) -> None:
        """init_relax_filters (domain_obj.f90:3285-3390) for the domain's
        boundary flags and halo width."""
        self._check(self.lib.icar_adv_set_relax_filter(self.h, int(west), int(east), int(north),
                                                        int(south), int(halo)))

    def set_forcing(self, name: str, forcing_hi: np.ndarray, dqdt: np.ndarray) -> None:
        """Uploads a species' boundary forcing and its tendency per second."""
        f = np.ascontiguousarray(forcing_hi, dtype=np.float32)
        t = np.ascontiguousarray(dqdt, dtype=np.float32)
        if f.shape != self.domain.shape or t.shape != self.domain.shape:
            raise ValueError("forcing arrays need the mass-grid shape")
        self._check(self.lib.icar_adv_set_forcing(self.h, self.names.index(name), _ptr(f),
                                                   _ptr(t)))

    def apply_forcing(self, dt: float) -> None:
        """apply_forcing (domain_obj.f90:3493-3503) of every forced species."""
        self._check(self.lib.icar_adv_apply_forcing(self.h, ctypes.c_double(dt)))

    # -- the step's tail and output ------------------------------------------
    def enforce_limits(self) -> None:
        """domain%enforce_limits (domain_obj.f90:3092-3150) on the device
        species: values below zero become zero."""
        self._check(self.lib.icar_adv_enforce_limits(self.h))

    def snapshot(self, species: bool = True, fields: bool = False) -> None:
        """Starts copying the current species (and fields) out for output or
        a restart file; the copy overlaps the following steps.
        ``snapshot_wait`` puts the values into the bound host arrays."""
        mask = (_lib.MASK_ALL_FIELDS if fields else 0) | (_lib.MASK_ALL_SPECIES if species else 0)
        self._check(self.lib.icar_adv_snapshot(self.h, mask))

    def snapshot_wait(self) -> None:
        self._check(self.lib.icar_adv_snapshot_wait(self.h))

    # -- halo exchange -------------------------------------------------------
    def halo_bytes(self, side: int) -> int:
        """Bytes of the strip side `side` (_lib.NORTH/SOUTH/EAST/WEST) sends."""
        n = ctypes.c_ulonglong()
        self._check(self.lib.icar_adv_halo_bytes(self.h, side, ctypes.byref(n)))
        return n.value

    def halo_pack(self, side: int, dev_ptr: int, nbytes: int) -> None:
        """The strip inside the tile next to `side` into device memory at
        dev_ptr (what a send to that neighbour carries)."""
        self._check(self.lib.icar_adv_halo_pack(self.h, side, ctypes.c_void_p(dev_ptr), nbytes))

    def halo_unpack(self, side: int, dev_ptr: int, nbytes: int) -> None:
        """Device memory at dev_ptr into the halo on `side`."""
        self._check(self.lib.icar_adv_halo_unpack(self.h, side, ctypes.c_void_p(dev_ptr), nbytes))

    def halo_exchange(self, side: int, to: "Advection") -> None:
        """This domain's strip on `side` into `to`'s halo on the opposite side
        (both on one device)."""
        self._check(self.lib.icar_adv_halo_exchange(self.h, side, to.h))

    def set_graph(self, enable: bool) -> None:
        """Record the run() loop in a HIP graph and replay it."""
        self._check(self.lib.icar_adv_set_graph(self.h, int(enable)))

    def profile(self, enable: bool) -> None:
        self._check(self.lib.icar_adv_profile(self.h, int(enable)))

    def launch_config(self) -> tuple:
        """(corr_cfg, ksplit) the fused path chose on its first step."""
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.icar_adv_launch_config(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def kernel_stats(self) -> dict:
        out = {}
        for kid in range(self.lib.icar_adv_kernel_count()):
            ms = ctypes.c_double()
            n = ctypes.c_longlong()
            b = ctypes.c_double()
            self._check(self.lib.icar_adv_kernel_stats(self.h, kid, ctypes.byref(ms),
                                                       ctypes.byref(n), ctypes.byref(b)))
            if n.value:
                out[self.lib.icar_adv_kernel_name(kid).decode()] = {
                    "total_ms": ms.value, "launches": n.value, "alg_bytes": b.value}
        return out


class Ensemble:
    """Independent MPDATA domains (one Advection each, kernel_variant 1)
    advanced together: one launch per pass over every (domain, tile,
    species), the step loop optionally recorded in a HIP graph."""

    def __init__(self, members: list):
        self.lib = _lib.load()
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        e = ctypes.c_void_p()
        _lib.check_ens(self.lib.icar_adv_ensemble_create(len(self.members), arr, ctypes.byref(e)))
        self.e = e

    def _check(self, rc):
        _lib.check_ens(rc, self.e)

    def run(self, dts, nsteps: int, bc_x: int = 0, bc_y: int = 0, graph: bool = False) -> None:
        dt = np.ascontiguousarray(dts, dtype=np.float32)
        assert dt.shape == (len(self.members),)
        self._check(self.lib.icar_adv_ensemble_run(self.e, dt.ctypes.data_as(_fp), nsteps, bc_x,
                                                    bc_y, int(graph)))

    def sync(self) -> None:
        self._check(self.lib.icar_adv_ensemble_sync(self.e))

    def profile(self, enable: bool) -> None:
        self._check(self.lib.icar_adv_ensemble_profile(self.e, int(enable)))

    def launch_config(self) -> tuple:
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.icar_adv_ensemble_launch_config(self.e, ctypes.byref(a),
                                                             ctypes.byref(b)))
        return a.value, b.value

    def kernel_stats(self) -> dict:
        out = {}
        for kid in range(self.lib.icar_adv_kernel_count()):
            ms, n, b = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
            self._check(self.lib.icar_adv_ensemble_kernel_stats(self.e, kid, ctypes.byref(ms),
                                                                ctypes.byref(n), ctypes.byref(b)))
            if n.value:
                out[self.lib.icar_adv_kernel_name(kid).decode()] = {
                    "total_ms": ms.value, "launches": n.value, "alg_bytes": b.value}
        return out

    def close(self) -> None:
        if getattr(self, "e", None) is not None and self.e.value:
            self.lib.icar_adv_ensemble_destroy(self.e)
            self.e = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# the name of the MPDATA-only first version of this class
MpdataAdvection = Advection


def adv_var_request(options: Options) -> None:
    """advection_driver.f90:46-57 -> adv_std_var_request (advect.f90:94-108):
    both schemes advect water vapour and potential temperature."""
    options.advect_vars(["qv", "theta"])


def adv_init(domain: Domain, options: Options) -> Advection:
    """advection_driver.f90:26-44: creates the device state (mpdata_init or
    adv_std_init) and uploads the domain; returns it and keeps it on the
    domain for ``advect``."""
    if options.physics.advection in (kADV_MPDATA, kADV_STD):
        domain._mpdata = Advection(domain, options)
        return domain._mpdata
    raise ValueError(f"unknown advection scheme {options.physics.advection}")


def balance_uvw(domain: Domain, options: Options) -> None:
    """wind.f90:88-164 (update=.false.) for a domain whose device state
    adv_init created: w is rebalanced on the device from the device u, v,
    density and metric (time_step.f90:324 runs it before advect when
    advect_density is set)."""
    st = getattr(domain, "_mpdata", None)
    if st is None:
        raise RuntimeError("adv_init(domain, options) has not been called")
    st.balance_w()


def compute_dt(domain: Domain, options: Options, dz_levels) -> float:
    """update_dt's compute_dt (time_step.f90:213-246) on the device winds,
    with options%time_options%cfl_reduction_factor and cfl_strictness."""
    st = getattr(domain, "_mpdata", None)
    if st is None:
        raise RuntimeError("adv_init(domain, options) has not been called")
    return st.compute_dt(options.time_options.cfl_reduction_factor,
                         options.time_options.cfl_strictness, dz_levels)


def advect(domain: Domain, options: Options, dt: float, sync_host: bool = True) -> None:
    """advection_driver.f90:59-126: the winds once (mpdata_compute_wind or
    adv_std_compute_wind), then every advected species in registry order, on
    the device.  adv=1 takes the three RK3 calls or one Euler call of
    adv_std_advect3d; under RK3 MPDATA does nothing (:112-115).  With
    ``sync_host`` the host species arrays are updated in place on return, as
    the reference's are."""
    if options.physics.advection not in (kADV_MPDATA, kADV_STD):
        raise ValueError(f"unknown advection scheme {options.physics.advection}")
    st = getattr(domain, "_mpdata", None)
    if st is None:
        raise RuntimeError("adv_init(domain, options) has not been called")
    st.advect(float(np.float32(dt)))
    if sync_host:
        st.download()
