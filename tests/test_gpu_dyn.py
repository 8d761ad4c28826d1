"""The wind balance (balance_uvw) and the time-step bound (compute_dt) on the
device, through the C-ABI, against the FP32 oracle bit for bit."""
import numpy as np
import pytest

from hicar_amd import _lib, synthetic
from hicar_amd.advection import MpdataAdvection, adv_init, balance_uvw, compute_dt

pytestmark = pytest.mark.gpu


def windy_case(seed, nx=70, ny=45, nz=12):
    case = synthetic.config1(nx=nx, ny=ny, nz=nz, steps=1)
    d = case.domain
    r = np.random.default_rng(seed)
    d.v[...] = r.uniform(-4, 4, d.v.shape)
    d.u[...] += r.uniform(-1, 1, d.u.shape).astype(np.float32)
    return case


def cols(d):
    """Balanced columns its-1..ite+1 x jts-1..jte+1, 0-based."""
    return (d.its - 1 - d.ims, d.ite + 1 - d.ims, d.jts - 1 - d.jms, d.jte + 1 - d.jms)


@pytest.mark.parametrize("adv", [True, False])
def test_balance_w_matches_oracle(oracle32, adv):
    case = windy_case(11)
    d = case.domain
    case.options.parameters.advect_density = adv
    ref = oracle32.balance_w(d, adv, *cols(d))
    st = MpdataAdvection(d, case.options)
    try:
        st.balance_w()
        st.download(species=False, fields=True)
    finally:
        st.close()
    np.testing.assert_array_equal(d.w, ref)
    assert np.abs(ref).max() > 0
    # the balanced field has no mass-flux divergence in the balanced columns
    full = oracle32.calc_divergence(d, adv, w=d.w, horz_only=False)
    i_s, i_e, j_s, j_e = cols(d)
    horz = oracle32.calc_divergence(d, adv)
    assert np.abs(full[j_s:j_e + 1, :, i_s:i_e + 1]).max() < 1e-4 * np.abs(horz).max()


@pytest.mark.parametrize("strictness", [1, 2, 3, 4, 5])
def test_compute_dt_matches_oracle(oracle32, strictness):
    case = windy_case(12)
    d = case.domain
    d.w[...] = oracle32.balance_w(d, True, *cols(d))
    dzl = synthetic.dz_levels(d.nz).astype(np.float32)
    tile = (d.its - d.ims, d.ite - d.ims, d.jts - d.jms, d.jte - d.jms)
    ref = np.float32(oracle32.compute_dt(d, dzl, 0.9, strictness, *tile))
    st = MpdataAdvection(d, case.options)
    try:
        if ref < 0.1:  # the reference stops there; the library returns an error with dt
            with pytest.raises(_lib.IcarAdvError) as e:
                st.compute_dt(0.9, strictness, dzl)
            assert e.value.code == _lib.ECOURANT and "too small" in str(e.value)
        else:
            got = st.compute_dt(0.9, strictness, dzl)
            assert np.float32(got) == ref, (got, ref)
    finally:
        st.close()


def test_host_mirror_balance_then_dt_then_advect(oracle32):
    """The order of a time step: balance_uvw before advect when
    advect_density is set (time_step.f90:324), dt from the balanced winds
    (update_dt), then advect; each step equals the oracle's."""
    case = windy_case(13)
    d, o = case.domain, case.options
    o.kernel_variant = 1
    dzl = synthetic.dz_levels(d.nz).astype(np.float32)
    w_ref = oracle32.balance_w(d, True, *cols(d))
    tile = (d.its - d.ims, d.ite - d.ims, d.jts - d.jms, d.jte - d.jms)
    init = {n: a.copy() for n, a in d.species.items()}
    st = adv_init(d, o)
    try:
        balance_uvw(d, o)
        dt = compute_dt(d, o, dzl)
        d2 = synthetic.config1(nx=d.nx, ny=d.ny, nz=d.nz, steps=1).domain
        for n in d.FIELDS:
            getattr(d2, n)[...] = getattr(d, n)
        d2.w[...] = w_ref
        assert np.float32(dt) == np.float32(oracle32.compute_dt(d2, dzl, 0.9, 3, *tile))
        st.advect(dt)
        st.download()
        ref = oracle32.advect(d2, o, dt, init)
        for n in ref:
            np.testing.assert_array_equal(d.species[n], ref[n], err_msg=n)
    finally:
        st.close()
