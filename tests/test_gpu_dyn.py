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


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("flags", [(1, 1, 1, 1), (1, 0, 0, 0), (0, 1, 0, 1), (0, 0, 1, 0)])
def test_lateral_forcing_matches_oracle(oracle32, variant, flags):
    """init_relax_filters + apply_forcing on the device copies: two forced
    species and one unforced, two forcing updates, bit for bit the oracle's
    (and the species the fused path keeps in its alternate buffers too)."""
    case = windy_case(14, nx=40, ny=33, nz=7)
    d, o = case.domain, case.options
    r = np.random.default_rng(15)
    o.vars_to_advect = ["qv", "qc", "qtr"]
    d.set_var_lists(o, {n: r.uniform(0.5, 2.0, d.shape) for n in o.vars_to_advect})
    o.kernel_variant = variant
    forcing = {n: (r.uniform(0.5, 2.0, d.shape).astype(np.float32),
                   r.uniform(-1e-4, 1e-4, d.shape).astype(np.float32)) for n in ("qv", "qtr")}
    filt = oracle32.relax_filter(d.nx, d.ny, 2, *flags)
    ref = {n: a.copy() for n, a in d.species.items()}
    fhi = {n: f.copy() for n, (f, _) in forcing.items()}
    st = MpdataAdvection(d, o)
    try:
        if variant == 1:  # the fused path keeps alternating home buffers
            st.advect(case.dt)
            st.download()
            ref = {n: a.copy() for n, a in d.species.items()}
        st.set_relax_filter(*flags, halo=2)
        for n, (f, t) in forcing.items():
            st.set_forcing(n, f, t)
        for dt in (37.5, 12.25):
            st.apply_forcing(dt)
            for n, (_, t) in forcing.items():
                ref[n], fhi[n] = oracle32.apply_forcing(ref[n], fhi[n], t, filt, dt)
        st.download()
    finally:
        st.close()
    for n in ref:
        np.testing.assert_array_equal(d.species[n], ref[n], err_msg=n)
    assert not np.array_equal(ref["qv"], ref["qc"])


def test_forcing_needs_fp32_storage():
    case = windy_case(16, nx=20, ny=12, nz=5)
    d, o = case.domain, case.options
    o.kernel_variant = 1
    d.storage["qv"] = (_lib.STORE_BF16, 0)
    st = MpdataAdvection(d, o)
    try:
        with pytest.raises(_lib.IcarAdvError):
            st.set_forcing("qv", d.species["qv"], np.zeros(d.shape, np.float32))
        with pytest.raises(_lib.IcarAdvError):  # no filter yet
            st.apply_forcing(10.0)
    finally:
        st.close()


@pytest.mark.parametrize("steps", [1, 2])
def test_lateral_forcing_after_standard_steps_matches_oracle(oracle32, steps):
    """The standard scheme swaps every species with its next-state buffer at
    each step; forcing after one or two steps (either buffer current) acts on
    the current state, and the next step advances the forced state: bit for
    bit the oracle doing the same."""
    from hicar_amd.advection import Advection
    case = synthetic.config_std(nx=36, ny=30, nz=7, steps=1, nspecies=3, cfl=1.0)
    d, o = case.domain, case.options
    r = np.random.default_rng(17)
    names = list(d.species)
    forcing = {n: (r.uniform(0.5, 2.0, d.shape).astype(np.float32),
                   r.uniform(-1e-4, 1e-4, d.shape).astype(np.float32)) for n in names[:2]}
    flags = (1, 1, 0, 1)
    filt = oracle32.relax_filter(d.nx, d.ny, d.halo, *flags)
    ref = {n: a.copy() for n, a in d.species.items()}
    for _ in range(steps):
        ref = oracle32.std_advect(d, o, case.dt, ref)
    fhi = {n: f.copy() for n, (f, _) in forcing.items()}
    for n, (_, t) in forcing.items():
        ref[n], fhi[n] = oracle32.apply_forcing(ref[n], fhi[n], t, filt, 20.0)
    ref = oracle32.std_advect(d, o, case.dt, ref)
    st = Advection(d, o)
    try:
        for _ in range(steps):
            st.advect(case.dt)
        st.set_relax_filter(*flags, halo=d.halo)
        for n, (f, t) in forcing.items():
            st.set_forcing(n, f, t)
        st.apply_forcing(20.0)
        st.advect(case.dt)
        st.download()
    finally:
        st.close()
    for n in ref:
        np.testing.assert_array_equal(d.species[n], ref[n], err_msg=n)
