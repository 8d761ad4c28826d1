"""Parity of the fused, species-batched path (kernel_variant 1) with the CPU
oracle and with the exact per-species kernels.

The fused kernels keep the reference's operation order, IEEE division and no
contraction, so for FP32 storage they reproduce the strict FP32 oracle bit
for bit (values; the sign of zero is not compared) over every order, FCT
setting, tile edge and species count exercised here.  This matters: the
MPDATA step on rough fields is sensitive to rounding order (the oracle
compiled with FMA contraction, as the reference's own -O2 -mfma build
contracts, moves theta of a 130x70x30 crop of the workload by ~6e-4 in 5
steps), so a reordered build could not meet the 1e-5 parity bar.
"""
import numpy as np
import pytest

from hicar_amd import synthetic
from hicar_amd.advection import MpdataAdvection

from test_gpu_parity import fresh_species, gpu_run, random_domain

pytestmark = pytest.mark.gpu


def assert_same(got, ref):
    for n in ref:
        np.testing.assert_array_equal(got[n], ref[n], err_msg=n)


@pytest.mark.parametrize("order,fct,dens", [(1, False, True), (2, False, True), (2, True, True),
                                            (3, True, True), (2, True, False), (4, True, True),
                                            (3, False, False)])
def test_batched_matches_oracle_random(oracle32, order, fct, dens):
    # 131 x 37: several tiles in x and y with partial edge tiles
    case = random_domain(131, 37, 9, 200 + order, order, fct, dens, nspec=3)
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, 0, 0, fresh_species(case))
    assert_same(gpu_run(case, variant=1), ref)


def test_batched_c1_100_steps(oracle32):
    case = synthetic.config1()
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                       fresh_species(case))
    assert_same(gpu_run(case, variant=1), ref)


def test_batched_c1_fct_100_steps(oracle32):
    case = synthetic.config1()
    case.options.adv_options.MPDATA_FCT = True
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                       fresh_species(case))
    assert_same(gpu_run(case, variant=1), ref)


def test_batched_c2_crop_50_steps(oracle32):
    """A 130x70x30 crop of the 10-species workload, order 3 with FCT, 50
    steps: bitwise equal to the oracle, and no species goes negative (the
    hydrometeors start 60 % zero)."""
    case = synthetic.config2(nx=130, ny=70, nz=30, steps=50)
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                       fresh_species(case))
    got = gpu_run(case, variant=1)
    assert_same(got, ref)
    for n, q in got.items():
        assert q.min() >= 0.0, (n, float(q.min()))


def test_batched_step_function_scenario(oracle32):
    case = synthetic.test_mpdata_case()
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                       fresh_species(case))
    got = gpu_run(case, variant=1)
    assert_same(got, ref)
    q = got["qtr"]
    assert q.max() <= 2.0 + 1e-6 and q.min() >= 1.0 - 1e-6


def test_batched_matches_exact_full_width():
    """Full 1024-wide rows (17 tiles + a partial one), 10 species, order 3:
    the batched path equals the exact per-species kernels."""
    case = synthetic.config2(nx=1024, ny=40, nz=30, steps=5)
    init = fresh_species(case)
    a = gpu_run(case, variant=0)
    for n, v in init.items():
        case.domain.species[n][...] = v
    assert_same(gpu_run(case, variant=1), a)


def test_batched_closed_box_mass():
    """Zero flux through the ring and the top: the FP64 budget of every
    species is conserved to FP32 rounding by the fused passes."""
    case = synthetic.config2(nx=70, ny=50, nz=20, steps=20)
    d = case.domain
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


@pytest.fixture
def launch_env(monkeypatch):
    """Fixes the corrective kernel's launch configuration through the
    environment the library reads when a handle builds its batch."""
    def set_cfg(cfg=None, ksplit=None):
        for k, v in (("ICAR_ADV_CORR_CFG", cfg), ("ICAR_ADV_KSPLIT", ksplit)):
            if v is None:
                monkeypatch.delenv(k, raising=False)
            else:
                monkeypatch.setenv(k, str(v))
    return set_cfg


@pytest.mark.parametrize("ksplit,order,fct", [(2, 3, True), (3, 2, True), (5, 3, True),
                                              (4, 2, False)])
def test_column_chunks_match_oracle(oracle32, launch_env, ksplit, order, fct):
    """The column march split into chunks (each started three levels early
    and storing only its own levels) gives the oracle's bits; nz = 17 makes
    the last chunk short."""
    launch_env(cfg=1, ksplit=ksplit)
    case = random_domain(131, 37, 17, 300 + ksplit, order, fct, True, nspec=2)
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, 0, 0, fresh_species(case))
    case.options.kernel_variant = 1
    st = MpdataAdvection(case.domain, case.options)
    try:
        st.run(case.dt, case.steps, 0, 0)
        st.download()
        cfg, ks = st.launch_config()
    finally:
        st.close()
    assert cfg == 1 and ks == min(ksplit, -(-17 // 4))
    assert_same({n: case.domain.species[n] for n in ref}, ref)


@pytest.mark.parametrize("cfg", [0, 2, 3, 5, 7])
def test_launch_configurations_match_oracle(oracle32, launch_env, cfg):
    """Every corrective-kernel configuration, the 60x32 tiles (cfg 5, 7)
    included, on a domain whose 37 rows leave partial tiles of either
    height."""
    launch_env(cfg=cfg, ksplit=1)
    case = random_domain(131, 37, 9, 400 + cfg, 3, True, True, nspec=2)
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, 0, 0, fresh_species(case))
    assert_same(gpu_run(case, variant=1), ref)


def test_autotuned_small_domain_matches_oracle(oracle32, launch_env):
    """The 256x64x20 ridge domain with one species leaves the GPU with 20
    column tiles: the autotuner times the launch configurations and column
    chunkings on the first step, and whatever it picks stays bit-exact."""
    launch_env()
    case = synthetic.config1(steps=30)
    case.options.adv_options.MPDATA_FCT = True
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                       fresh_species(case))
    case.options.kernel_variant = 1
    st = MpdataAdvection(case.domain, case.options)
    try:
        assert st.launch_config() == (-1, -1)
        st.run(case.dt, case.steps, case.bc_x, case.bc_y)
        st.download()
        cfg, ks = st.launch_config()
    finally:
        st.close()
    assert 0 <= cfg <= 7 and 1 <= ks <= 5
    print(f"\n[autotune] 256x64x20, 1 species: corr_cfg {cfg}, column chunks {ks}")
    assert_same({"qv": case.domain.species["qv"]}, ref)


@pytest.mark.parametrize("fct", [True, False])
def test_fast_division_outside_its_range_matches_oracle(oracle32, fct):
    """Species spanning 1e-30..1e25 (numerators below 2^-78 and above 2^64),
    and, without FCT, species of both signs (divisors r + l + 1e-10 near zero
    or negative) give the oracle's values bit for bit.  With the opt-in
    three-operation quotient (ICAR_ADV_FAST_DIV=1) such operands are outside
    its range and the waves holding them take the IEEE division; the same
    case runs under that setting in tests/fastdiv_check.py."""
    case = random_domain(131, 37, 9, 77, 3, fct, True, nspec=4)
    d = case.domain
    r = np.random.default_rng(78)
    names = list(d.species)
    d.species[names[0]][...] *= np.float32(1e-30)
    d.species[names[1]][...] *= np.float32(1e25)
    d.species[names[2]][...] = (10.0 ** r.uniform(-38, 20, d.shape)).astype(np.float32) * \
        (r.random(d.shape) > 0.3)
    if not fct:
        d.species[names[3]][...] -= np.float32(0.5)
    ref = oracle32.run(d, case.options, case.dt, case.steps, 0, 0, fresh_species(case))
    assert_same(gpu_run(case, variant=1), ref)
