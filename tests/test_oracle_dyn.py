"""Pins the CPU restatement of the wind balance and the time-step bound
(oracle/dyn_oracle.c) to what the mathematics fixes:

  calc_divergence   closed forms for linear winds and a linearly varying
                    density, including the one-sided density at the domain
                    edges (wind.f90:222-227)
  calc_w            after balancing, the full 3-D mass-flux divergence
                    (calc_divergence with horz_only=.false.) vanishes, with
                    and without advect_density; closed form of w for a
                    uniform horizontal divergence; zero outside the columns
                    it is asked to balance
  compute_dt        closed forms of every cfl_strictness for uniform winds,
                    and an independent FP64 NumPy evaluation of the tile
                    maximum (the w of the level below, zoffset) for varying
                    winds
Every test runs on the FP32 and the FP64 build.
"""
import numpy as np
import pytest

from hicar_amd import synthetic
from hicar_amd.domain import Domain


@pytest.fixture(params=["f32", "f64"])
def orc(request, oracle32, oracle64):
    return oracle32 if request.param == "f32" else oracle64


def tol(orc):
    return 1e-12 if orc.dtype == np.float64 else 2e-5


def plain_domain(nx=9, ny=7, nz=5, dx=500.0):
    d = Domain(nx, ny, nz, dx)
    for n in ("density", "jacobian", "jacobian_u", "jacobian_v", "jacobian_w"):
        getattr(d, n)[...] = 1.0
    d.advection_dz[...] = 2.0
    return d


# --------------------------------------------------------------------------
# calc_divergence (wind.f90:201-246)
# --------------------------------------------------------------------------
def test_divergence_of_linear_winds(orc):
    d = plain_domain()
    xf = np.arange(d.nx + 1, dtype=np.float64)
    yf = np.arange(d.ny + 1, dtype=np.float64)
    d.u[...] = 3.0 * xf[None, None, :]          # du/di = 3
    d.v[...] = -1.25 * yf[:, None, None]        # dv/dj = -1.25
    for adv in (False, True):
        div = orc.calc_divergence(d, adv)
        np.testing.assert_allclose(div, (3.0 - 1.25) / d.dx, rtol=tol(orc))


def test_divergence_density_weights_and_edges(orc):
    """u constant, density linear in i: the face density is the average of
    the two cells inside and the linear extrapolation 1.5 rho0 - 0.5 rho1 at
    the two edges, so d(rho u)/dx = u * drho/di / dx in every cell."""
    d = plain_domain()
    ii = np.arange(d.nx, dtype=np.float64)
    d.density[...] = 1.0 + 0.125 * ii[None, None, :]
    d.u[...] = 4.0
    d.v[...] = 0.0
    div = orc.calc_divergence(d, True)
    np.testing.assert_allclose(div, 4.0 * 0.125 / d.dx, rtol=tol(orc))
    # without advect_density the density does not enter
    np.testing.assert_allclose(orc.calc_divergence(d, False), 0.0, atol=tol(orc))
    # the same along y
    jj = np.arange(d.ny, dtype=np.float64)
    d.density[...] = 2.0 - 0.0625 * jj[:, None, None]
    d.u[...] = 0.0
    d.v[...] = -2.0
    np.testing.assert_allclose(orc.calc_divergence(d, True), -2.0 * -0.0625 / d.dx, rtol=tol(orc))


# --------------------------------------------------------------------------
# calc_w (wind.f90:166-199) through balance_uvw
# --------------------------------------------------------------------------
@pytest.mark.parametrize("adv", [False, True])
def test_balanced_winds_have_no_divergence(orc, adv):
    case = synthetic.config1(nx=24, ny=9, nz=8, steps=1)
    d = case.domain
    d.v[...] = np.random.default_rng(5).uniform(-3, 3, d.v.shape)
    w = orc.balance_w(d, adv)
    full = orc.calc_divergence(d, adv, w=w, horz_only=False)
    horz = orc.calc_divergence(d, adv)
    assert np.abs(horz).max() > 1e-5
    assert np.abs(full).max() <= 50 * tol(orc) * np.abs(horz).max()


def test_w_closed_form_and_columns(orc):
    """A uniform horizontal divergence D with jacobian_w = rho = 1 and dz = 2
    gives w(k) = -D * dz * (k + 1); columns outside i_s..i_e, j_s..j_e are 0."""
    d = plain_domain()
    xf = np.arange(d.nx + 1, dtype=np.float64)
    d.u[...] = 0.5 * xf[None, None, :]
    d.v[...] = 0.0
    D = 0.5 / d.dx
    for adv in (False, True):
        w = orc.balance_w(d, adv, 2, 6, 1, 4)
        kk = np.arange(d.nz, dtype=np.float64)
        exp = -D * 2.0 * (kk + 1)
        inside = w[1:5, :, 2:7]
        np.testing.assert_allclose(inside, np.broadcast_to(exp[None, :, None], inside.shape),
                                   rtol=tol(orc))
        mask = np.ones(w.shape, bool)
        mask[1:5, :, 2:7] = False
        assert np.all(w[mask] == 0)


# --------------------------------------------------------------------------
# compute_dt (time_step.f90:51-168)
# --------------------------------------------------------------------------
def test_compute_dt_closed_forms(orc):
    d = plain_domain(dx=1000.0)
    U, V, W = 12.0, -5.0, 0.75
    d.u[...] = U
    d.v[...] = V
    d.w[...] = W
    dzl = np.array([40.0, 50.0, 60.0, 80.0, 100.0])
    cfl = 0.9
    sqrt3 = np.float64(np.sqrt(np.float32(3.0)) * np.float32(1.001)) if orc.dtype == np.float32 \
        else np.sqrt(3.0) * 1.001
    m3 = max(U / d.dx, abs(V) / d.dx, W / dzl.min())  # sum over directions is not taken
    m1 = max(U, abs(V), W)                               # m/s, as the reference compares
    exp = {1: cfl / (m1 * sqrt3), 3: cfl / m3, 4: cfl / (m3 * sqrt3),
           5: cfl / (U + abs(V) + W), 2: cfl / max(m1, m3 * 0.577350269)}
    for s, e in exp.items():
        assert orc.compute_dt(d, dzl, cfl, s) == pytest.approx(e, rel=tol(orc) * 10), s


def test_compute_dt_tile_maximum(orc):
    """Varying winds: the strictness-3 maximum over the tile of
    max(|u| on the two x faces / dx, |v| on the two y faces / dx,
    |w| of the level and the level below / dz), against NumPy in FP64; cells
    outside the tile do not count, the whole arrays do for strictness 1."""
    rng = np.random.default_rng(9)
    d = plain_domain(nx=11, ny=8, nz=6, dx=800.0)
    d.u[...] = rng.uniform(-15, 15, d.u.shape)
    d.v[...] = rng.uniform(-15, 15, d.v.shape)
    d.w[...] = rng.uniform(-2, 2, d.w.shape)
    dzl = rng.uniform(30, 90, d.nz)
    its, ite, jts, jte = 2, 8, 1, 6
    u = np.abs(d.u.astype(orc.dtype).astype(np.float64))
    v = np.abs(d.v.astype(orc.dtype).astype(np.float64))
    w = np.abs(d.w.astype(orc.dtype).astype(np.float64))
    dz = dzl.astype(orc.dtype).astype(np.float64)
    a = np.maximum(u[:, :, :-1], u[:, :, 1:]) / d.dx
    b = np.maximum(v[:-1], v[1:]) / d.dx
    wb = np.concatenate([w[:, :1], w[:, :-1]], axis=1)  # the level below (itself at k = 0)
    c = np.maximum(w, wb) / dz[None, :, None]
    cur = np.maximum(np.maximum(a, b), c)[jts:jte + 1, :, its:ite + 1]
    got = orc.compute_dt(d, dzl, 0.9, 3, its, ite, jts, jte)
    assert got == pytest.approx(0.9 / cur.max(), rel=tol(orc) * 10)
    # a larger wind outside the tile changes strictness 1 only
    d.u[0, 0, 0] = 500.0
    assert orc.compute_dt(d, dzl, 0.9, 3, its, ite, jts, jte) == got
    assert orc.compute_dt(d, dzl, 0.9, 1, its, ite, jts, jte) == pytest.approx(
        0.9 / (500.0 * np.sqrt(3.0) * 1.001), rel=1e-6)
