"""ctypes binding of ``libicar_adv.so`` (declared in ``include/icar_adv.h``).

The library must have been built (``python -m hicar_amd.build`` or
``__graft_entry__.build()``); there is no fallback: loading fails loudly when
it is missing.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libicar_adv.so")

# enums of icar_adv.h
FIELD_U, FIELD_V, FIELD_W, FIELD_DENSITY, FIELD_JACOBIAN = 0, 1, 2, 3, 4
FIELD_JACOBIAN_U, FIELD_JACOBIAN_V, FIELD_JACOBIAN_W, FIELD_ADVECTION_DZ = 5, 6, 7, 8
NFIELDS = 9
FIELD_NAMES = ("u", "v", "w", "density", "jacobian", "jacobian_u", "jacobian_v",
               "jacobian_w", "advection_dz")
STORE_F32, STORE_BF16, STORE_F16, STORE_BF16_SR, STORE_F16_SR = 0, 1, 2, 3, 4
STORE_NAMES = {0: "f32", 1: "bf16", 2: "f16", 3: "bf16-sr", 4: "f16-sr"}
MAX_SPECIES = 48
MASK_ALL_FIELDS = 0x1FF
MASK_ALL_SPECIES = 0xFFFFFFFFFFFF0000
EINVAL, ESHAPE, EUNBOUND, ERANGE, ENOMEM, ECOURANT = -1, -2, -3, -4, -5, -6
ADV_STD, ADV_MPDATA = 1, 2
NORTH, SOUTH, EAST, WEST = 0, 1, 2, 3


def mask_species(s: int) -> int:
    return 1 << (16 + s)


EXPORTS = (
    "icar_adv_create", "icar_adv_destroy", "icar_adv_bind_field", "icar_adv_bind_species",
    "icar_adv_upload", "icar_adv_download", "icar_adv_advect", "icar_adv_mass_budget",
    "icar_adv_max_courant", "icar_adv_last_error", "icar_adv_sync", "icar_adv_apply_bc",
    "icar_adv_run", "icar_adv_profile", "icar_adv_kernel_count", "icar_adv_kernel_name",
    "icar_adv_kernel_stats", "icar_adv_set_graph", "icar_adv_ensemble_create",
    "icar_adv_ensemble_destroy", "icar_adv_ensemble_run", "icar_adv_ensemble_sync",
    "icar_adv_ensemble_last_error", "icar_adv_ensemble_profile", "icar_adv_ensemble_kernel_stats",
    "icar_adv_launch_config", "icar_adv_ensemble_launch_config", "icar_adv_balance_w",
    "icar_adv_compute_dt", "icar_adv_set_relax_filter", "icar_adv_set_forcing",
    "icar_adv_apply_forcing", "icar_adv_halo_bytes", "icar_adv_halo_pack",
    "icar_adv_halo_unpack", "icar_adv_halo_exchange", "icar_adv_enforce_limits",
    "icar_adv_snapshot", "icar_adv_snapshot_wait", "icar_adv_divergence",
    "icar_adv_mass_budget_noalloc", "icar_adv_verify_division", "icar_adv_set_compensation",
)


class Grid(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("ims", "ime", "kms", "kme", "jms", "jme", "its", "ite", "jts", "jte")] + \
               [("dx", ctypes.c_float)]


class Opts(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("mpdata_order", "mpdata_fct", "advect_density", "debug", "rk3",
                 "kernel_variant", "advection", "h_order", "v_order", "flux_corr")]


class IcarAdvError(RuntimeError):
    """A non-zero return code of the C-ABI, with the library's message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"icar_adv error {code}: {msg}")
        self.code = code


_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first (python -m hicar_amd.build)")
    lib = ctypes.CDLL(LIB_PATH)
    vp, ip, fp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_float)
    i3 = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    sig = {
        "icar_adv_create": ([ctypes.POINTER(Grid), ctypes.POINTER(Opts), ctypes.POINTER(vp)], ip),
        "icar_adv_destroy": ([vp], ip),
        "icar_adv_bind_field": ([vp, ip, fp, i3, i3], ip),
        "icar_adv_bind_species": ([vp, ip, fp, ip, ip], ip),
        "icar_adv_upload": ([vp, ctypes.c_ulonglong], ip),
        "icar_adv_download": ([vp, ctypes.c_ulonglong], ip),
        "icar_adv_advect": ([vp, ctypes.c_float], ip),
        "icar_adv_mass_budget": ([vp, dp], ip),
        "icar_adv_max_courant": ([vp, dp], ip),
        "icar_adv_divergence": ([vp, dp], ip),
        "icar_adv_set_compensation": ([vp, ip], ip),
        "icar_adv_mass_budget_noalloc": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp,
                                          vp, vp, vp], ip),
        "icar_adv_verify_division": ([ctypes.c_uint, ctypes.c_uint,
                                      ctypes.POINTER(ctypes.c_ulonglong),
                                      ctypes.POINTER(ctypes.c_ulonglong)], ip),
        "icar_adv_last_error": ([vp], ctypes.c_char_p),
        "icar_adv_sync": ([vp], ip),
        "icar_adv_apply_bc": ([vp, ip, ip], ip),
        "icar_adv_run": ([vp, ctypes.c_float, ip, ip, ip], ip),
        "icar_adv_profile": ([vp, ip], ip),
        "icar_adv_kernel_count": ([], ip),
        "icar_adv_kernel_name": ([ip], ctypes.c_char_p),
        "icar_adv_kernel_stats": ([vp, ip, dp, ctypes.POINTER(ctypes.c_longlong), dp], ip),
        "icar_adv_set_graph": ([vp, ip], ip),
        "icar_adv_ensemble_create": ([ip, ctypes.POINTER(vp), ctypes.POINTER(vp)], ip),
        "icar_adv_ensemble_destroy": ([vp], ip),
        "icar_adv_ensemble_run": ([vp, fp, ip, ip, ip, ip], ip),
        "icar_adv_ensemble_sync": ([vp], ip),
        "icar_adv_ensemble_last_error": ([vp], ctypes.c_char_p),
        "icar_adv_ensemble_profile": ([vp, ip], ip),
        "icar_adv_ensemble_kernel_stats": ([vp, ip, dp, ctypes.POINTER(ctypes.c_longlong), dp], ip),
        "icar_adv_launch_config": ([vp, i3, i3], ip),
        "icar_adv_balance_w": ([vp], ip),
        "icar_adv_compute_dt": ([vp, ctypes.c_float, ip, fp, fp], ip),
        "icar_adv_set_relax_filter": ([vp, ip, ip, ip, ip, ip], ip),
        "icar_adv_set_forcing": ([vp, ip, fp, fp], ip),
        "icar_adv_apply_forcing": ([vp, ctypes.c_double], ip),
        "icar_adv_ensemble_launch_config": ([vp, i3, i3], ip),
        "icar_adv_halo_bytes": ([vp, ip, ctypes.POINTER(ctypes.c_ulonglong)], ip),
        "icar_adv_halo_pack": ([vp, ip, vp, ctypes.c_ulonglong], ip),
        "icar_adv_halo_unpack": ([vp, ip, vp, ctypes.c_ulonglong], ip),
        "icar_adv_halo_exchange": ([vp, ip, vp], ip),
        "icar_adv_enforce_limits": ([vp], ip),
        "icar_adv_snapshot": ([vp, ctypes.c_ulonglong], ip),
        "icar_adv_snapshot_wait": ([vp], ip),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(rc: int, handle=None) -> None:
    if rc != 0:
        msg = load().icar_adv_last_error(handle)
        raise IcarAdvError(rc, msg.decode() if msg else "")


def check_ens(rc: int, ens=None) -> None:
    if rc != 0:
        msg = load().icar_adv_ensemble_last_error(ens)
        raise IcarAdvError(rc, msg.decode() if msg else "")
