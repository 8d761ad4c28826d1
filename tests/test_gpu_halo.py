"""Halo exchange of the advected species between subdomains (domain%
halo_exchange before advect, time_step.f90:362-366; the strips of
domain_obj.f90:573-650 and exchangeable_obj.f90:240-392), on one GPU:

  - data movement: a side sends the strip inside the tile as wide as its
    halo, over the tile's range along the side, and the neighbour writes it
    into its halo on the opposite side; nothing else changes (corners are
    not exchanged), through icar_adv_halo_exchange and through
    icar_adv_halo_pack / _unpack with a device buffer;
  - a domain split into a west and an east image that exchange before each
    step, advanced by the library, equals the FP32 oracle doing the same
    exchange (numpy slices) and steps on the two images, bit for bit, for
    the standard scheme and for MPDATA.
"""
import copy
import ctypes

import numpy as np
import pytest

from hicar_amd import _lib, synthetic
from hicar_amd.advection import Advection
from hicar_amd.domain import Domain

pytestmark = pytest.mark.gpu

FIELDS_X1 = ("u", "jacobian_u")  # nx+1 faces per row


def image(case, i0, nxi):
    """The subdomain of case's domain whose memory columns are
    [i0, i0 + nxi) of the global arrays, with the same halo."""
    d = case.domain
    sub = Domain(nxi, d.ny, d.nz, d.dx, halo=d.halo)
    for n in d.FIELDS:
        a = getattr(d, n)
        w = nxi + 1 if n in FIELDS_X1 else nxi
        getattr(sub, n)[...] = a[..., i0:i0 + w]
    sub.set_var_lists(case.options, {n: a[..., i0:i0 + nxi].copy() for n, a in d.species.items()})
    return sub


def split(case, mid):
    d = case.domain
    h = d.halo
    west = image(case, 0, mid + h)
    east = image(case, mid - h, d.nx - mid + h)
    return west, east


def tile(d):
    return (d.its - d.ims, d.ite - d.ims, d.jts - d.jms, d.jte - d.jms)


def exchange_x(west, east, qw, qe):
    """The west image's east strip into the east image's west halo and back,
    on the tile rows (numpy restatement of the two strips)."""
    h = west.halo
    wi0, wi1, j0, j1 = tile(west)
    ei0, _, _, _ = tile(east)
    rows = slice(j0, j1 + 1)
    for n in qw:
        send_w = qw[n][rows, :, wi1 - h + 1:wi1 + 1].copy()
        send_e = qe[n][rows, :, ei0:ei0 + h].copy()
        qe[n][rows, :, 0:ei0] = send_w
        qw[n][rows, :, wi1 + 1:wi1 + 1 + h] = send_e


def random_case(nx=40, ny=30, nz=6, names=("qv", "qc", "qr")):
    case = synthetic.config_std(nx=nx, ny=ny, nz=nz, steps=1, nspecies=len(names), h_order=5,
                                v_order=3, flux_corr=0, rk3=True, cfl=1.0)
    r = np.random.default_rng(21)
    for a in case.domain.species.values():
        a[...] = r.uniform(0, 1, a.shape)
    return case


def test_exchange_moves_the_strips_and_nothing_else():
    case = random_case()
    west, east = split(case, 20)
    r = np.random.default_rng(3)
    for dom in (west, east):  # distinct contents everywhere, halos included
        for a in dom.species.values():
            a[...] = r.uniform(0, 1, a.shape)
    qw = {n: a.copy() for n, a in west.species.items()}
    qe = {n: a.copy() for n, a in east.species.items()}
    sw, se = Advection(west, case.options), Advection(east, case.options)
    try:
        sw.halo_exchange(_lib.EAST, se)
        se.halo_exchange(_lib.WEST, sw)
        sw.sync()
        se.sync()
        sw.download()
        se.download()
    finally:
        sw.close()
        se.close()
    exchange_x(west, east, qw, qe)
    for n in qw:
        np.testing.assert_array_equal(west.species[n], qw[n], err_msg=n)
        np.testing.assert_array_equal(east.species[n], qe[n], err_msg=n)
    # the corners of the halo were left alone
    j0 = tile(west)[2]
    assert np.array_equal(west.species["qv"][:j0], qw["qv"][:j0])


class DeviceBuffer:
    """Device memory from the HIP runtime (plumbing for the test)."""

    def __init__(self, nbytes):
        import ctypes
        import os
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        self.hip = ctypes.CDLL(path)
        self.ptr = ctypes.c_void_p()
        rc = self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes))
        assert rc == 0, rc

    def close(self):
        if self.ptr.value:
            self.hip.hipFree(self.ptr)
            self.ptr.value = None


def test_pack_unpack_all_four_sides_through_a_device_buffer():
    case = random_case(nx=34, ny=34)
    d = case.domain
    src = Domain(d.nx, d.ny, d.nz, d.dx, halo=d.halo)
    dst = Domain(d.nx, d.ny, d.nz, d.dx, halo=d.halo)
    for n in d.FIELDS:
        getattr(src, n)[...] = getattr(d, n)
        getattr(dst, n)[...] = getattr(d, n)
    r = np.random.default_rng(4)
    src.set_var_lists(case.options, {n: r.uniform(0, 1, d.shape) for n in d.species})
    dst.set_var_lists(case.options, {n: r.uniform(0, 1, d.shape) for n in d.species})
    want = {n: a.copy() for n, a in dst.species.items()}
    i0, i1, j0, j1 = tile(d)
    h = d.halo
    ti, tj = slice(i0, i1 + 1), slice(j0, j1 + 1)
    for n, a in src.species.items():  # the reference's strips, side by side
        want[n][j0 - h:j0, :, ti] = a[j1 - h + 1:j1 + 1, :, ti]   # north strip -> south halo
        want[n][j1 + 1:j1 + 1 + h, :, ti] = a[j0:j0 + h, :, ti]   # south strip -> north halo
        want[n][tj, :, i0 - h:i0] = a[tj, :, i1 - h + 1:i1 + 1]   # east strip -> west halo
        want[n][tj, :, i1 + 1:i1 + 1 + h] = a[tj, :, i0:i0 + h]   # west strip -> east halo
    ss, sd = Advection(src, case.options), Advection(dst, case.options)
    try:
        for side, opp in ((_lib.NORTH, _lib.SOUTH), (_lib.SOUTH, _lib.NORTH),
                          (_lib.EAST, _lib.WEST), (_lib.WEST, _lib.EAST)):
            nb = ss.halo_bytes(side)
            assert nb == sd.halo_bytes(opp)
            buf = DeviceBuffer(nb)
            try:
                ss.halo_pack(side, buf.ptr.value, nb)
                ss.sync()
                sd.halo_unpack(opp, buf.ptr.value, nb)
                sd.sync()
                with pytest.raises(_lib.IcarAdvError):  # a buffer one element short
                    ss.halo_pack(side, buf.ptr.value, nb - 4)
            finally:
                buf.close()
        sd.download()
    finally:
        ss.close()
        sd.close()
    for n in want:
        np.testing.assert_array_equal(dst.species[n], want[n], err_msg=n)


@pytest.mark.parametrize("scheme", ["std", "mpdata"])
def test_two_images_exchange_and_advance_like_the_oracle(oracle32, scheme):
    if scheme == "std":
        case = synthetic.config_std(nx=44, ny=30, nz=8, steps=3, nspecies=3, cfl=1.0)
    else:
        case = synthetic.config2(nx=44, ny=30, nz=8, steps=3, nspecies=3)
        case.options.kernel_variant = 1
    west, east = split(case, 22)
    qw = {n: a.copy() for n, a in west.species.items()}
    qe = {n: a.copy() for n, a in east.species.items()}
    sw, se = Advection(west, case.options), Advection(east, case.options)
    try:
        for _ in range(case.steps):
            sw.halo_exchange(_lib.EAST, se)
            se.halo_exchange(_lib.WEST, sw)
            sw.advect(case.dt)
            se.advect(case.dt)
        sw.download()
        se.download()
    finally:
        sw.close()
        se.close()
    for _ in range(case.steps):
        exchange_x(west, east, qw, qe)
        if scheme == "std":
            qw = oracle32.std_advect(west, case.options, case.dt, qw)
            qe = oracle32.std_advect(east, case.options, case.dt, qe)
        else:
            qw = oracle32.advect(west, case.options, case.dt, qw)
            qe = oracle32.advect(east, case.options, case.dt, qe)
    for n in qw:
        np.testing.assert_array_equal(west.species[n], qw[n], err_msg=n)
        np.testing.assert_array_equal(east.species[n], qe[n], err_msg=n)


def mpdata_images(storage_w, e_w, storage_e, e_e):
    """West and east images of an MPDATA domain (fused path) whose species
    qc is stored in 16 bits with the given formats and 2^e scales."""
    case = synthetic.config2(nx=44, ny=30, nz=6, steps=1, nspecies=3)
    case.options.kernel_variant = 1
    west, east = split(case, 22)
    west.storage["qc"] = (storage_w, e_w)
    east.storage["qc"] = (storage_e, e_e)
    return case, west, east


def test_16bit_species_exchange_their_stored_values():
    d0 = synthetic.config2(nx=44, ny=30, nz=6, steps=1, nspecies=3).domain
    e = synthetic.pow2_scale(d0.species["qc"], _lib.STORE_BF16)
    case, west, east = mpdata_images(_lib.STORE_BF16, e, _lib.STORE_BF16_SR, e)
    sw, se = Advection(west, case.options), Advection(east, case.options)
    try:
        sw.download()  # the values as stored (qc rounded to BF16)
        se.download()
        qw = {n: a.copy() for n, a in west.species.items()}
        qe = {n: a.copy() for n, a in east.species.items()}
        sw.halo_exchange(_lib.EAST, se)
        se.halo_exchange(_lib.WEST, sw)
        sw.sync()
        se.sync()
        sw.download()
        se.download()
    finally:
        sw.close()
        se.close()
    exchange_x(west, east, qw, qe)
    for n in qw:
        np.testing.assert_array_equal(west.species[n], qw[n], err_msg=n)
        np.testing.assert_array_equal(east.species[n], qe[n], err_msg=n)


@pytest.mark.parametrize("other", ["scale", "format"])
def test_exchange_refuses_species_stored_differently(other):
    d0 = synthetic.config2(nx=44, ny=30, nz=6, steps=1, nspecies=3).domain
    e = synthetic.pow2_scale(d0.species["qc"], _lib.STORE_BF16)
    if other == "scale":
        case, west, east = mpdata_images(_lib.STORE_BF16, e, _lib.STORE_BF16, e + 1)
    else:
        case, west, east = mpdata_images(_lib.STORE_BF16, e, _lib.STORE_F16, e)
    sw, se = Advection(west, case.options), Advection(east, case.options)
    try:
        with pytest.raises(_lib.IcarAdvError) as err:
            sw.halo_exchange(_lib.EAST, se)
        assert err.value.code == _lib.ESHAPE and "stored differently" in str(err.value)
    finally:
        sw.close()
        se.close()


def test_exchange_refuses_different_species_of_equal_count():
    """The strips are paired by species: sides binding species {0, 1} and
    {0, 2} hold the same number of strips, and the exchange refuses them
    rather than write species 1 into species 2's halo."""
    case = synthetic.config2(nx=44, ny=30, nz=6, steps=1, nspecies=2)
    case.options.kernel_variant = 1
    west, east = split(case, 22)
    first = list(east.species)[0]
    o1 = copy.deepcopy(case.options)
    o1.vars_to_advect = [first]
    east.set_var_lists(o1, {first: east.species[first]})
    sw, se = Advection(west, case.options), Advection(east, o1)
    extra = np.ones(east.shape, np.float32)
    try:
        se._check(se.lib.icar_adv_bind_species(
            se.h, 2, extra.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _lib.STORE_F32, 0))
        with pytest.raises(_lib.IcarAdvError) as err:
            sw.halo_exchange(_lib.EAST, se)
        assert err.value.code == _lib.ESHAPE and "same species" in str(err.value)
    finally:
        sw.close()
        se.close()
