"""Parity of the fused, species-batched path (kernel_variant 1) with the CPU
oracle and with the exact kernels.

The fused kernels use reciprocals and FMA contraction, so they agree with the
reference order to rounding, not bit for bit.  The bar is the north star's:
max|q - q_ref| / max|q_ref| <= 1e-5 per field after the stated number of
steps (100 for the 256x64x20 ridge case), mass conserved to FP32 rounding in
a closed box, and the FCT bound of the reference's step-function scenario.

Where a field is itself sensitive to rounding order, a non-bitwise build of
the reference cannot meet 1e-5 either: the oracle compiled with FMA
contraction (the rounding of the reference's own -O2 -mfma build) moves theta
of the 1024x1024 workload by ~6e-4 after 5 steps, the growth starting at the
un-upwinded top flux.  For those fields the bar is that spread: the fused
path must be no further from the strict oracle than 2x the FMA build is.
"""
import numpy as np
import pytest

from hicar_amd import synthetic
from hicar_amd.advection import MpdataAdvection

from test_gpu_parity import fresh_species, gpu_run, norm_err, random_domain

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.parametrize("order,fct,dens", [(1, False, True), (2, False, True), (2, True, True),
                                            (3, True, True), (2, True, False), (4, True, True),
                                            (3, False, False)])
def test_fast_matches_oracle_random(oracle32, order, fct, dens):
    # 131 x 37: several tiles in x and y with partial edge tiles
    case = random_domain(131, 37, 9, 200 + order, order, fct, dens, nspec=3)
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, 0, 0, fresh_species(case))
    got = gpu_run(case, variant=1)
    for n in ref:
        assert norm_err(got[n], ref[n]) <= TOL, (n, norm_err(got[n], ref[n]))


def test_fast_c1_100_steps(oracle32):
    case = synthetic.config1()
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                       fresh_species(case))
    got = gpu_run(case, variant=1)
    assert norm_err(got["qv"], ref["qv"]) <= TOL


def test_fast_c1_fct_100_steps(oracle32):
    case = synthetic.config1()
    case.options.adv_options.MPDATA_FCT = True
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                       fresh_species(case))
    got = gpu_run(case, variant=1)
    assert norm_err(got["qv"], ref["qv"]) <= TOL


def fma_spread(oracle32, oracle32_fma, case, ref):
    alt = oracle32_fma.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                           fresh_species(case))
    return {n: norm_err(alt[n], ref[n]) for n in ref}


def test_fast_c2_crop_50_steps(oracle32, oracle32_fma):
    case = synthetic.config2(nx=130, ny=70, nz=30, steps=50)
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                       fresh_species(case))
    spread = fma_spread(oracle32, oracle32_fma, case, ref)
    got = gpu_run(case, variant=1)
    for n in ref:
        e = norm_err(got[n], ref[n])
        assert e <= max(TOL, 2 * spread[n]), (n, e, spread[n])


def test_fast_step_function_scenario():
    case = synthetic.test_mpdata_case()
    got = gpu_run(case, variant=1)
    q = got["qtr"]
    assert q.max() <= 2.0 + 1e-6 and q.min() >= 1.0 - 1e-6
    assert abs(q[:, 1, 1].mean() - 1.5) < 1e-3


def test_fast_matches_exact_full_width(oracle32, oracle32_fma):
    """Full 1024-wide rows (17 tiles + a partial one), 10 species, order 3,
    against the exact kernels (bit-identical to the strict oracle)."""
    case = synthetic.config2(nx=1024, ny=40, nz=30, steps=5)
    init = fresh_species(case)
    a = gpu_run(case, variant=0)
    alt = oracle32_fma.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                           init)
    spread = {n: norm_err(alt[n], a[n]) for n in a}
    for n, v in init.items():
        case.domain.species[n][...] = v
    b = gpu_run(case, variant=1)
    for n in a:
        e = norm_err(b[n], a[n])
        assert e <= max(TOL, 2 * spread[n]), (n, e, spread[n])


def test_fast_closed_box_mass():
    """Zero flux through the ring and the top: the FP64 budget of every
    species is conserved to FP32 rounding by the fused passes."""
    case = synthetic.config2(nx=70, ny=50, nz=20, steps=20)
    d = case.domain
    d.u[:, :, :2] = 0
    d.u[:, :, -2:] = 0
    d.v[:2] = 0
    d.v[-2:] = 0
    # zero divergence column by column so that nothing leaves through the top
    d.w[...] = 0
    d.u[...] = 0
    d.v[...] = 0
    rng = np.random.default_rng(3)
    psi = np.zeros((d.ny + 1, d.nz, d.nx + 1))
    psi[2:-2, :, 2:-2] = rng.uniform(-200, 200, psi[2:-2, :, 2:-2].shape)
    d.u[:, :, 1:-1] = (psi[1:, :, 1:-1] - psi[:-1, :, 1:-1]) / d.dx * 10
    d.v[1:-1] = -(psi[1:-1, :, 1:] - psi[1:-1, :, :-1]) / d.dx * 10
    d.density[...] = 1.0
    d.jacobian_u[...] = 1.0
    d.jacobian_v[...] = 1.0
    case.options.kernel_variant = 1
    st = MpdataAdvection(d, case.options)
    try:
        m0 = st.mass_budget()
        st.run(case.dt, case.steps)
        m1 = st.mass_budget()
    finally:
        st.close()
    rel = np.abs(m1 - m0) / np.abs(m0)
    assert rel.max() < 2e-6, rel
