"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

The "exact" kernels (kernel_variant 0) keep the reference's IEEE FP32
operation order, so they are compared element by element with the FP32
oracle (bit-identical values); the parity bar for longer runs is
max|q - q_ref| / max|q_ref| <= 1e-5 per field.
"""
import numpy as np
import pytest

from hicar_amd import synthetic
from hicar_amd.advection import MpdataAdvection, adv_init, advect
from hicar_amd.domain import Domain, Options

pytestmark = pytest.mark.gpu

def assert_bits(got, ref, name):
    np.testing.assert_array_equal(np.asarray(got, np.float32).view(np.uint32),
                                  np.asarray(ref, np.float32).view(np.uint32), err_msg=name)


def gpu_run(case, steps=None, variant=0):
    case.options.kernel_variant = variant
    st = MpdataAdvection(case.domain, case.options)
    try:
        st.run(case.dt, case.steps if steps is None else steps, case.bc_x, case.bc_y)
        st.download()
        return {n: case.domain.species[n].copy() for n in st.names}
    finally:
        st.close()


def fresh_species(case):
    return {n: a.copy() for n, a in case.domain.species.items()}


def random_domain(nx, ny, nz, seed, order, fct, advect_density, nspec=2):
    r = np.random.default_rng(seed)
    d = Domain(nx, ny, nz, 1000.0)
    d.u[...] = r.uniform(-12, 12, d.u.shape)
    d.v[...] = r.uniform(-12, 12, d.v.shape)
    d.w[...] = r.uniform(-0.5, 0.5, d.w.shape)
    d.density[...] = 0.8 + 0.4 * r.random(d.shape)
    d.jacobian[...] = 0.7 + 0.3 * r.random(d.shape)
    d.jacobian_u[...] = 0.7 + 0.3 * r.random(d.jacobian_u.shape)
    d.jacobian_v[...] = 0.7 + 0.3 * r.random(d.jacobian_v.shape)
    d.jacobian_w[...] = 0.7 + 0.3 * r.random(d.shape)
    d.advection_dz[...] = (50 * 1.15 ** np.arange(nz))[None, :, None]
    o = Options()
    o.adv_options.mpdata_order = order
    o.adv_options.MPDATA_FCT = fct
    o.parameters.advect_density = advect_density
    names = [f"s{i}" for i in range(nspec)]
    o.vars_to_advect = names
    vals = {n: (r.random(d.shape) + 0.01) * (r.random(d.shape) > 0.5) for n in names}
    d.set_var_lists(o, vals)
    return synthetic.Case("random", d, o, 20.0, 3, 0, 0)


@pytest.mark.parametrize("order,fct,dens", [(0, True, True), (1, False, True), (2, False, True),
                                            (2, True, True), (3, True, True), (2, True, False),
                                            (4, True, False)])
def test_exact_matches_oracle_random(oracle32, order, fct, dens):
    """The per-species kernels against the FP32 oracle, bit for bit; order 0
    runs no pass of the order loop (adv_mpdata.f90:367) and leaves q as is."""
    case = random_domain(37, 21, 9, 100 + order, order, fct, dens)
    init = fresh_species(case)
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, 0, 0, fresh_species(case))
    got = gpu_run(case)
    for n in ref:
        assert_bits(got[n], ref[n], n)
        if order == 0:
            assert_bits(got[n], init[n], n)


def test_exact_c1_100_steps(oracle32):
    case = synthetic.config1()
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                       fresh_species(case))
    got = gpu_run(case)
    assert_bits(got["qv"], ref["qv"], "qv")


def test_exact_c2_crop(oracle32):
    case = synthetic.config2(nx=96, ny=80, nz=30, steps=10)
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                       fresh_species(case))
    got = gpu_run(case)
    for n in ref:
        assert_bits(got[n], ref[n], n)


def test_reference_step_function_scenario_gpu(oracle32):
    case = synthetic.test_mpdata_case()
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                       fresh_species(case))
    got = gpu_run(case)
    np.testing.assert_array_equal(got["qtr"], ref["qtr"])
    assert got["qtr"].max() <= 2.0 + 1e-6 and got["qtr"].min() >= 1.0 - 1e-6


def test_rk3_is_noop_and_host_interface():
    case = synthetic.config1(nx=24, ny=9, nz=6, steps=1)
    d, o = case.domain, case.options
    q0 = d.species["qv"].copy()
    o.time_options.RK3 = True
    adv_init(d, o)
    advect(d, o, case.dt)
    np.testing.assert_array_equal(d.species["qv"], q0)
    d._mpdata.close()
    o.time_options.RK3 = False
    adv_init(d, o)
    advect(d, o, case.dt)
    assert np.abs(d.species["qv"] - q0).max() > 0
    d._mpdata.close()


def test_mass_budget_and_courant(oracle32):
    case = synthetic.config1(nx=40, ny=12, nz=8, steps=1)
    d = case.domain
    st = MpdataAdvection(d, case.options)
    try:
        st.advect(case.dt)
        m = st.mass_budget()
        st.download()
        exp = oracle32.mass(oracle32.arr(d.species["qv"]), oracle32.arr(d.jacobian),
                            oracle32.arr(d.density), oracle32.arr(d.advection_dz))
        assert m[0] == pytest.approx(exp, rel=1e-12)
        c = st.max_courant()
        # the FP32 Courant numbers of the step, reduced exactly (a max needs no
        # rounding); W/dz as an FP32 quotient, as the step divides
        U, V, W = oracle32.compute_wind(d, case.dt, True)
        Wd = W / oracle32.arr(d.advection_dz)
        assert c[0] == float(np.abs(U[1:-1, :, 1:]).max())
        assert c[2] == float(np.abs(Wd[1:-1, :, 1:-1]).max())
        assert c[1] == float(np.abs(V[1:, :, 1:-1]).max())
        assert 0 < max(c) <= 0.5 * 1.3  # dt at Courant 0.5 of the physical winds
    finally:
        st.close()


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("balanced", [True, False])
def test_divergence_check_through_the_abi(oracle32, variant, balanced, capfd):
    """test_divergence (adv_mpdata.f90:445-475): the count of interior cells
    with |div| > 1e-3 and the largest |div| of the step's Courant numbers,
    equal to the oracle's on the same FP32 Courant numbers; with debug set
    advect also prints the warning when the count is not zero."""
    if balanced:  # w from discrete continuity: the reference's check stays quiet
        case = synthetic.config1(nx=40, ny=12, nz=8, steps=1)
    else:  # random winds: most cells exceed 1e-3
        case = random_domain(30, 20, 7, 31, 2, True, True)
    d, o = case.domain, case.options
    o.kernel_variant = variant
    o.parameters.debug = True
    st = MpdataAdvection(d, o)
    try:
        with pytest.raises(Exception, match="advect has not been called"):
            st.divergence()
        st.advect(case.dt)
        cnt, mx = st.divergence()
    finally:
        st.close()
    U, V, W = oracle32.compute_wind(d, case.dt, o.parameters.advect_density)
    n_ref, mx_ref = oracle32.divergence(U, V, W, oracle32.arr(d.advection_dz))
    assert (cnt, mx) == (n_ref, mx_ref)
    err = capfd.readouterr().err
    if balanced:
        assert n_ref == 0 and mx_ref < 1e-4 and "not balanced" not in err
    else:
        assert n_ref > (d.nx - 2) * (d.ny - 2) * d.nz // 2
        assert f"{n_ref} cells" in err


def test_divergence_check_refused_for_the_standard_scheme():
    std = synthetic.config_std(nx=20, ny=16, nz=6, steps=1, nspecies=1, cfl=1.0)
    st = MpdataAdvection(std.domain, std.options)
    try:
        st.advect(std.dt)
        with pytest.raises(Exception, match="belongs to MPDATA"):
            st.divergence()
    finally:
        st.close()


def test_upload_download_roundtrip_and_shape_errors():
    case = synthetic.config1(nx=20, ny=7, nz=5, steps=1)
    d = case.domain
    st = MpdataAdvection(d, case.options)
    try:
        q = d.species["qv"]
        orig = q.copy()
        q[...] = -1.0
        st.download()
        np.testing.assert_array_equal(q, orig)
    finally:
        st.close()
    from hicar_amd import _lib
    import ctypes
    lib = _lib.load()
    g = _lib.Grid(1, 20, 1, 5, 1, 7, 3, 18, 3, 5, 1000.0)
    o = _lib.Opts(2, 1, 0, 0, 0, 0)
    h = ctypes.c_void_p()
    _lib.check(lib.icar_adv_create(ctypes.byref(g), ctypes.byref(o), ctypes.byref(h)))
    try:
        lb = (ctypes.c_int * 3)(1, 1, 1)
        ub = (ctypes.c_int * 3)(20, 5, 7)  # u needs ime+1
        arr = np.zeros((7, 5, 21), np.float32)
        rc = lib.icar_adv_bind_field(h, _lib.FIELD_U,
                                     arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), lb, ub)
        assert rc == -2 and b"do not match" in lib.icar_adv_last_error(h)
        assert lib.icar_adv_advect(h, ctypes.c_float(1.0)) == -3  # fields unbound
    finally:
        lib.icar_adv_destroy(h)


def test_exact_is_bit_identical_to_oracle(oracle32):
    """The exact kernels reproduce the FP32 oracle bit for bit (values; the
    sign of zero is not compared)."""
    case = random_domain(29, 17, 11, 7, 3, True, True, nspec=3)
    ref = oracle32.run(case.domain, case.options, case.dt, 5, 0, 0, fresh_species(case))
    got = gpu_run(case, steps=5)
    for n in ref:
        np.testing.assert_array_equal(got[n], ref[n])
    case = synthetic.config1(steps=20)
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                       fresh_species(case))
    got = gpu_run(case)
    np.testing.assert_array_equal(got["qv"], ref["qv"])
