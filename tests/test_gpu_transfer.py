"""Host transfers of the species and fields (icar_adv_upload/_download).
Transfers of 8 MB and more go through the library's pinned staging in 32 MB
chunks, the DMA of one chunk overlapping the CPU copy of the other, so these
tests use arrays of several chunks and sizes that are not multiples of a
chunk:
  - a round trip returns every value bit for bit, FP32 species and fields;
  - what the upload put on the device is what the host held: the device's
    FP64 mass budget equals numpy's sum of the same products to rounding,
    after the host arrays were overwritten (so it saw the device copy);
  - 16-bit species go through the FP32 staging and come back as the FP32
    rounding of the packed values, the same bits as a small transfer gives.
"""
import numpy as np
import pytest

from hicar_amd import _lib, synthetic
from hicar_amd.advection import MpdataAdvection
from hicar_amd.domain import Domain, Options

pytestmark = pytest.mark.gpu


def big_case(nx=700, ny=530, nz=31, names=("qv", "qc", "qr", "qtr")):
    """About 46 MB per species: 1.4 chunks, not a multiple of a chunk."""
    d = Domain(nx, ny, nz, 1000.0)
    r = np.random.default_rng(5)
    for n in d.FIELDS:
        a = getattr(d, n)
        a[...] = r.uniform(0.5, 1.5, a.shape)
    o = Options()
    o.adv_options.mpdata_order = 2
    o.vars_to_advect = list(names)
    d.set_var_lists(o, {n: r.uniform(0.0, 1.0, d.shape) for n in names})
    return d, o


def test_round_trip_is_bitwise_through_the_pinned_staging():
    d, o = big_case()
    want_sp = {n: a.copy() for n, a in d.species.items()}
    want_f = {n: getattr(d, n).copy() for n in d.FIELDS}
    st = MpdataAdvection(d, o)
    try:
        for a in d.species.values():
            a[...] = -7.0
        for n in d.FIELDS:
            getattr(d, n)[...] = -7.0
        st.download(species=True, fields=True)
        for n in want_sp:
            np.testing.assert_array_equal(d.species[n], want_sp[n], err_msg=n)
        for n in want_f:
            np.testing.assert_array_equal(getattr(d, n), want_f[n], err_msg=n)
        # a second upload of changed values and a second round trip
        r = np.random.default_rng(6)
        for n, a in d.species.items():
            a[...] = r.uniform(0.0, 2.0, a.shape)
            want_sp[n] = a.copy()
        st.upload(fields=False, species=True)
        for a in d.species.values():
            a[...] = 0.0
        st.download()
        for n in want_sp:
            np.testing.assert_array_equal(d.species[n], want_sp[n], err_msg=n)
    finally:
        st.close()


def test_uploaded_values_reach_the_device():
    d, o = big_case()
    q = {n: a.astype(np.float64) for n, a in d.species.items()}
    w = d.jacobian.astype(np.float64) * d.density * d.advection_dz
    inner = (slice(1, -1), slice(None), slice(1, -1))  # the cells advect updates
    want = np.array([(q[n] * w)[inner].sum() for n in d.species])
    st = MpdataAdvection(d, o)
    try:
        for a in d.species.values():
            a[...] = 0.0
        got = st.mass_budget()
    finally:
        st.close()
    np.testing.assert_allclose(got, want, rtol=1e-12)


@pytest.mark.parametrize("storage", [_lib.STORE_BF16, _lib.STORE_F16])
def test_16bit_species_through_the_staging_match_a_small_transfer(storage):
    d, o = big_case(names=("qc", "qr"))
    for n, a in d.species.items():
        d.storage[n] = (storage, synthetic.pow2_scale(a, storage))
    init = {n: a.copy() for n, a in d.species.items()}
    o.kernel_variant = 1
    st = MpdataAdvection(d, o)
    try:
        st.download()
    finally:
        st.close()
    # the same values through a domain small enough for the direct copies
    ny = 8
    s = Domain(d.nx, ny, d.nz, 1000.0)
    s.set_var_lists(o, {n: a[:ny] for n, a in init.items()})
    s.storage = dict(d.storage)
    st = MpdataAdvection(s, o)
    try:
        st.download()
    finally:
        st.close()
    for n in init:
        np.testing.assert_array_equal(d.species[n][:ny], s.species[n], err_msg=n)
        assert not np.array_equal(d.species[n], init[n])  # rounded to 16 bits
        np.testing.assert_allclose(d.species[n], init[n], rtol=2 ** -7)


def test_large_snapshot_holds_the_state_it_was_taken_at():
    """A snapshot of several chunks' worth of species (copied into the host
    arrays by the copy pool) against a download of the same state."""
    d, o = big_case()
    o.kernel_variant = 1
    st = MpdataAdvection(d, o)
    try:
        st.download()
        want = {n: a.copy() for n, a in d.species.items()}
        st.snapshot(species=True)
        for a in d.species.values():
            a[...] = -1.0
        st.snapshot_wait()
    finally:
        st.close()
    for n in want:
        np.testing.assert_array_equal(d.species[n], want[n], err_msg=n)
