"""Pins the CPU oracle (oracle/mpdata_oracle.c) to what the mathematics of
the reference algorithm fixes.  The reference (src/physics/adv_mpdata.f90,
adv_mpdata_FCT_core.f90) cannot be built here and its own MPDATA test
(src/tests/test_mpdata.f90) asserts nothing, so every oracle function is
checked against closed forms, invariants and special cases chosen so that a
dropped term, a wrong sign or index or a transposed operand fails one of them:

  compute_wind (a3)   closed forms of U/V/W incl. the top level and the
                      advect_density switch
  upwind (a5)         exact one-cell shifts at Courant +-1 in x, y, z; the
                      un-upwinded top flux; constant fields under
                      non-divergent flow; conservation in a closed box
  fluxes (a6)         1-D reduction to the textbook antidiffusive velocity
                      (Smolarkiewicz 1983) times 1/2; closed forms of all six
                      cross terms on linear fields; the row guards
  fct (a7)            the FCT bound property (result inside the local
                      min/max of q0 and q1) incl. the w-column ends; the
                      x/y boundary cells left unlimited (fin = fout = 0, so
                      beta = room/1e-15) against the w-column ends limited
                      by their one face, and a boundary cell at its own
                      minimum zeroing its outflow; the step-function
                      scenario of test_mpdata.f90:51-111
  advect3d (a4)       order 1 == upwind; 1-D order 2 == textbook MPDATA
                      with the 0.5 damping and the dz rescaling of w (FP64
                      NumPy restatement); order 3 == order 2 plus one pass
                      with q0 = q1 (the iord >= 3 quirk); x/y mirror and
                      x<->y transposition symmetry of the full FCT step
  advect (a1)         RK3 no-op; species independence; density shadowing
                      (Courant numbers with rho=1, divergence with density)
  divergence, mass    closed forms
Every test runs on the FP32 and the FP64 build.
"""
import numpy as np
import pytest

from hicar_amd.domain import Domain, Options
from hicar_amd import synthetic


@pytest.fixture(params=["f32", "f64"])
def orc(request, oracle32, oracle64):
    return oracle32 if request.param == "f32" else oracle64


def rnd(seed):
    return np.random.default_rng(seed)


def fields(orc, shape, G=1.0, dz=1.0):
    ny, nz, nx = shape
    G = orc.arr(np.broadcast_to(G, shape))
    dz = orc.arr(np.broadcast_to(dz, shape))
    rho = orc.arr(np.ones(shape))
    return G, dz, rho


def zeros(orc, shape):
    return orc.arr(np.zeros(shape))


# --------------------------------------------------------------------------
# compute_wind (adv_mpdata.f90:477-504)
# --------------------------------------------------------------------------
def test_compute_wind_closed_form(orc):
    nx, ny, nz = 5, 4, 6
    d = Domain(nx, ny, nz, 4.0)
    d.u[...] = 2.0
    d.v[...] = -1.0
    d.w[...] = 0.5
    # density varies along each axis so that a wrong neighbour shows
    ii = np.arange(nx)[None, None, :]
    kk = np.arange(nz)[None, :, None]
    jj = np.arange(ny)[:, None, None]
    d.density[...] = 1.0 + ii + 10 * kk + 100 * jj
    d.jacobian_u[...] = 0.5
    d.jacobian_v[...] = 0.25
    d.jacobian_w[...] = 2.0
    dt = 3.0
    U, V, W = orc.compute_wind(d, dt, advect_density=True)
    rho = d.density.astype(np.float64)
    eU = 2.0 * dt * (rho[:, :, 1:] + rho[:, :, :-1]) * 0.5 * 0.5 / 4.0
    eV = -1.0 * dt * (rho[1:] + rho[:-1]) * 0.5 * 0.25 / 4.0
    eW = 0.5 * dt * 2.0 * (rho[:, 1:] + rho[:, :-1]) * 0.5
    np.testing.assert_allclose(U[:, :, 1:], eU, rtol=1e-6)
    np.testing.assert_array_equal(U[:, :, 0], 0)
    np.testing.assert_allclose(V[1:], eV, rtol=1e-6)
    np.testing.assert_array_equal(V[0], 0)
    np.testing.assert_allclose(W[:, :-1], eW, rtol=1e-6)
    # model top: w*dt*jacobian_w*rho, no average (:498)
    np.testing.assert_allclose(W[:, -1], 0.5 * dt * 2.0 * rho[:, -1], rtol=1e-6)
    # advect_density off: rho = 1 (:489)
    U1, V1, W1 = orc.compute_wind(d, dt, advect_density=False)
    np.testing.assert_allclose(U1[:, :, 1:], 2.0 * dt * 0.5 / 4.0, rtol=1e-6)
    np.testing.assert_allclose(W1[:, -1], 0.5 * dt * 2.0, rtol=1e-6)
    np.testing.assert_allclose(W1[:, :-1], 0.5 * dt * 2.0, rtol=1e-6)


# --------------------------------------------------------------------------
# upwind_advection (adv_mpdata.f90:43-103)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("axis,sign", [(0, 1), (0, -1), (1, 1), (1, -1)])
def test_upwind_courant_one_is_exact_shift(orc, axis, sign):
    nx, ny, nz = 9, 8, 5
    shape = (ny, nz, nx)
    q = orc.arr(rnd(1).random(shape) + 0.5)
    U, V, W = zeros(orc, shape), zeros(orc, shape), zeros(orc, shape)
    if axis == 0:
        U[:, :, 1:] = sign
    else:
        V[1:] = sign
    G, dz, rho = fields(orc, shape)
    out = orc.upwind(q, U, V, W, rho, dz, G)
    inner = (slice(1, -1), slice(None), slice(1, -1))
    if axis == 0:
        exp = np.roll(q, sign, axis=2)
    else:
        exp = np.roll(q, sign, axis=0)
    np.testing.assert_array_equal(out[inner], exp[inner])
    # ring copied unchanged (:61-63)
    np.testing.assert_array_equal(out[0], q[0])
    np.testing.assert_array_equal(out[:, :, -1], q[:, :, -1])


def test_upwind_vertical_shift_and_top_quirk(orc):
    nx, ny, nz = 5, 5, 7
    shape = (ny, nz, nx)
    q = orc.arr(rnd(2).random(shape) + 0.5)
    U, V = zeros(orc, shape), zeros(orc, shape)
    G, dz, rho = fields(orc, shape)
    Wup = orc.arr(np.ones(shape))
    out = orc.upwind(q, U, V, Wup, rho, dz, G)
    c = (slice(1, -1), slice(None), slice(1, -1))
    o, qq = out[c], q[c]
    np.testing.assert_array_equal(o[:, 1:], qq[:, :-1])     # shift up, top included
    np.testing.assert_array_equal(o[:, 0], 0 * qq[:, 0])    # no inflow through the ground
    Wdn = orc.arr(-np.ones(shape))
    out = orc.upwind(q, U, V, Wdn, rho, dz, G)
    o = out[c]
    np.testing.assert_array_equal(o[:, 1:-1], qq[:, 2:])    # shift down
    # nothing leaves through the ground: the bottom cell keeps its value and
    # receives the one above
    np.testing.assert_array_equal(o[:, 0], qq[:, 0] + qq[:, 1])
    # top flux qin(kme)*w(kme) is not upwinded (:96-97): the top cell keeps
    # its value instead of taking the (unknown) value above it
    np.testing.assert_array_equal(o[:, -1], qq[:, -1])


def closed_box_flow(orc, nx, ny, nz, seed, amp=4):
    """Courant numbers with zero flux through every face of the outer ring and
    the model top, exactly divergence free in floating point (dyadic values):
    a horizontal streamfunction psi (per level) and an x-z one phi (per row)."""
    r = rnd(seed)
    U = np.zeros((ny, nz, nx))
    V = np.zeros((ny, nz, nx))
    W = np.zeros((ny, nz, nx))
    psi = np.zeros((ny + 1, nz, nx + 1))   # corners (j, k, i)
    psi[2:ny - 1, :, 2:nx - 1] = r.integers(-amp, amp + 1, (ny - 3, nz, nx - 3)) / 64.0
    U[:, :, 1:] = psi[1:ny + 1, :, 1:nx] - psi[0:ny, :, 1:nx]
    V[1:, :, :] = -(psi[1:ny, :, 1:nx + 1] - psi[1:ny, :, 0:nx])
    phi = np.zeros((ny, nz + 1, nx + 1))   # x-z corners: (j, k-face below, i)
    phi[1:ny - 1, 1:nz, 2:nx - 1] = r.integers(-amp, amp + 1, (ny - 2, nz - 1, nx - 3)) / 64.0
    U[:, :, 1:] += phi[:, 1:, 1:nx] - phi[:, :-1, 1:nx]
    W[:, :, :] = -(phi[:, 1:, 1:] - phi[:, 1:, :-1])
    U[:, :, 1] = 0
    U[:, :, nx - 1] = 0
    V[1] = 0
    V[ny - 1] = 0
    W[:, nz - 1] = 0
    return orc.arr(U), orc.arr(V), orc.arr(W)


def divergence64(U, V, W):
    ny, nz, nx = U.shape
    dU = U[:, :, 2:].astype(np.float64) - U[:, :, 1:-1]
    dV = V[2:].astype(np.float64) - V[1:-1]
    Wp = np.concatenate([np.zeros((ny, 1, nx)), W.astype(np.float64)], axis=1)
    dW = Wp[:, 1:] - Wp[:, :-1]
    return dU[1:-1] + dV[:, :, 1:-1] + dW[1:-1, :, 1:-1]


def test_closed_box_flow_is_divergence_free(orc):
    U, V, W = closed_box_flow(orc, 12, 10, 6, 3)
    assert np.abs(divergence64(U, V, W)).max() == 0.0
    assert np.abs(U).max() > 0 and np.abs(V).max() > 0 and np.abs(W).max() > 0


@pytest.mark.parametrize("order,fct", [(1, False), (2, False), (2, True), (3, True)])
def test_constant_field_stays_constant(orc, order, fct):
    nx, ny, nz = 12, 10, 6
    shape = (ny, nz, nx)
    U, V, W = closed_box_flow(orc, nx, ny, nz, 4)
    G = orc.arr(0.75 + 0.5 * rnd(5).random(shape))
    dz = orc.arr(np.ones(shape))
    q = orc.arr(np.full(shape, 1.5))
    out = orc.advect3d(q, G, orc.arr(np.ones(shape)), dz, U, V, W, order, fct)
    np.testing.assert_array_equal(out, q)


@pytest.mark.parametrize("order,fct", [(1, False), (2, False), (2, True), (3, True), (3, False)])
def test_closed_box_conserves_mass(orc, order, fct):
    nx, ny, nz = 12, 10, 6
    shape = (ny, nz, nx)
    U, V, W = closed_box_flow(orc, nx, ny, nz, 6)
    r = rnd(7)
    rho = orc.arr(0.8 + 0.4 * r.random(shape))
    jaco = orc.arr(0.9 + 0.2 * r.random(shape))
    dz = orc.arr(np.broadcast_to((50 * 1.15 ** np.arange(nz))[None, :, None], shape))
    q = orc.arr(r.random(shape) * (r.random(shape) > 0.4))
    m0 = orc.mass(q, jaco, rho, dz)
    for _ in range(5):
        q = orc.advect3d(q, rho, jaco, dz, U, V, W, order, fct)
    m1 = orc.mass(q, jaco, rho, dz)
    tol = 1e-5 if orc.dtype == np.float32 else 1e-13
    assert abs(m1 - m0) / m0 < tol
    # the field did move
    assert np.abs(q - orc.arr(r.random(shape))).max() > 0


def test_c1_constant_field_and_divergence(orc):
    """Synthetic C1 winds are non-divergent for the scheme: a constant field
    stays constant to rounding through the full advect."""
    case = synthetic.config1(nx=32, ny=8, nz=10, steps=1)
    d = case.domain
    U, V, W = orc.compute_wind(d, case.dt, True)
    n, mx = orc.divergence(U, V, W, orc.arr(d.advection_dz))
    assert n == 0 and mx < 1e-5
    spec = {"c": np.full(d.shape, 0.01)}
    for order, fct in [(2, False), (2, True), (3, True)]:
        case.options.adv_options.mpdata_order = order
        case.options.adv_options.MPDATA_FCT = fct
        out = orc.advect(d, case.options, case.dt, spec)["c"]
        # the FP32-stored winds are non-divergent to their own rounding
        assert np.abs(out / 0.01 - 1).max() < 1e-5


# --------------------------------------------------------------------------
# mpdata_fluxes (adv_mpdata.f90:105-252) + the 0.5 damping (:378-380)
# --------------------------------------------------------------------------
def scaled_fluxes(orc, q, U, V, W, G, dz):
    u2, v2, w2 = orc.fluxes(q, U, V, W / dz, G)
    return u2 * 0.5, v2 * 0.5, w2 * 0.5 * dz


def test_fluxes_1d_reduce_to_textbook(orc):
    """V = W = 0, G = 1: u2 = (|U| - U^2) (r - l) / (r + l + 1e-10), the
    antidiffusive velocity of Smolarkiewicz (1983), halved (:378)."""
    nx, ny, nz = 11, 6, 5
    shape = (ny, nz, nx)
    r = rnd(8)
    q = orc.arr(r.random(shape) + 0.1)
    U = orc.arr(r.uniform(-0.9, 0.9, shape))
    U[:, :, 0] = 0
    Z = zeros(orc, shape)
    G, dz, _ = fields(orc, shape)
    u2, v2, w2 = scaled_fluxes(orc, q, U, Z, Z, G, dz)
    q64, U64 = q.astype(np.float64), U.astype(np.float64)
    rr, ll = q64[:, :, 1:], q64[:, :, :-1]
    exp = 0.5 * (np.abs(U64[:, :, 1:]) - U64[:, :, 1:] ** 2) * (rr - ll) / (rr + ll + 1e-10)
    np.testing.assert_allclose(u2[:, :, 1:], exp, rtol=2e-5, atol=1e-7)
    assert np.abs(v2).max() == 0 and np.abs(w2).max() == 0


@pytest.mark.parametrize("face,cross", [("u", "v"), ("u", "w"), ("v", "u"), ("v", "w"),
                                        ("w", "u"), ("w", "v")])
def test_fluxes_cross_terms_closed_form(orc, face, cross):
    """Uniform Courant numbers a (normal) and b (transverse), G = 1, dz = 1,
    q linear along the transverse axis only: the first term vanishes (r = l)
    and the face value is the cross term alone,
        -0.5 * 0.5 * a * b * A / (G + G),
    A = (sum of the two cells ahead - the two behind)/(sum of the four),
    present only where the reference's guards allow it."""
    nx, ny, nz = 8, 7, 6
    shape = (ny, nz, nx)
    a, b = 0.375, -0.25
    axis = {"u": 2, "v": 0, "w": 1}
    ii = np.arange(nx)[None, None, :]
    jj = np.arange(ny)[:, None, None]
    kk = np.arange(nz)[None, :, None]
    coord = {"u": ii, "v": jj, "w": kk}[cross]
    q = orc.arr(np.broadcast_to(2.0 + 0.25 * coord, shape))
    comp = {"u": zeros(orc, shape), "v": zeros(orc, shape), "w": zeros(orc, shape)}
    comp[face][...] = a
    comp[cross][...] = b
    comp["u"][:, :, 0] = 0
    comp["v"][0] = 0
    G, dz, _ = fields(orc, shape)
    u2, v2, w2 = scaled_fluxes(orc, q, comp["u"], comp["v"], comp["w"], G, dz)
    got = {"u": u2, "v": v2, "w": w2}[face]
    # the transverse coordinate at the two cells "ahead" is c+1, "behind" c-1;
    # for the w-face cross terms the stencil is shifted: compute A directly
    q64 = q.astype(np.float64)

    def A_at(c):
        # ratio (sum ahead - sum behind) / (sum all) for two cells along the
        # transverse axis at c+1 and c-1, both identical along the face normal
        up = 2.0 + 0.25 * (c + 1)
        dn = 2.0 + 0.25 * (c - 1)
        return (2 * up - 2 * dn) / (2 * up + 2 * dn)

    if face == "w" and cross == "u":
        # WxU uses x-neighbours i+1 and i-1 at levels k and k+1 (:227-232)
        A = A_at(ii * np.ones(shape))
    elif face == "w" and cross == "v":
        A = A_at(jj * np.ones(shape))
    elif face == "u" and cross == "w":
        A = A_at(kk * np.ones(shape))
    elif face == "v" and cross == "w":
        A = A_at(kk * np.ones(shape))
    elif face == "u" and cross == "v":
        A = A_at(jj * np.ones(shape))
    else:  # v x u
        A = A_at(ii * np.ones(shape))
    exp = -0.5 * 0.5 * a * b * A / 2.0
    # where the reference computes the term (guards of :145,157,186-190,197,227-231,238)
    mask = np.zeros(shape, bool)
    ny_, nz_, nx_ = shape
    if face == "u":
        mask[:, :, 1:] = True
        if cross == "v":
            mask[0] = mask[-1] = False
        else:
            mask[:, 0] = mask[:, -1] = False
    elif face == "v":
        mask[1:] = True
        if cross == "u":
            mask[:, :, 0] = mask[:, :, -1] = False
        else:
            mask[:, 0] = mask[:, -1] = False
    else:
        mask[:, :-1] = True
        if cross == "u":
            mask[:, :, 0] = mask[:, :, -1] = False
        else:
            mask[0] = mask[-1] = False
    np.testing.assert_allclose(got[mask], exp[mask], rtol=2e-5)
    assert np.abs(got[~mask]).max() == 0.0
    assert np.abs(exp[mask]).min() > 0


# --------------------------------------------------------------------------
# FCT (adv_mpdata_FCT_core.f90:47-116)
# --------------------------------------------------------------------------
def flux1(l, r, U):
    return np.where(U > 0, U * l, U * r)


@pytest.mark.parametrize("is_w", [False, True])
def test_fct_keeps_result_inside_local_bounds(orc, is_w):
    r = rnd(9)
    for trial in range(200):
        n = int(r.integers(3, 14))
        q0 = r.random(n) * (r.random(n) > 0.3)
        q1 = np.clip(q0 + 0.3 * r.standard_normal(n), 0, None)
        U2 = r.uniform(-1.0, 1.0, n - 1)
        lim = orc.fct_core(q1, q0, U2, is_w).astype(np.float64)
        q0d, q1d = orc.arr(q0).astype(np.float64), orc.arr(q1).astype(np.float64)
        # limiting only shrinks each face towards zero, keeping its sign
        U2d = orc.arr(U2).astype(np.float64)
        assert np.all(np.abs(lim) <= np.abs(U2d) * (1 + 1e-6) + 1e-30)
        assert np.all(lim * U2d >= 0)
        F = flux1(q1d[:-1], q1d[1:], lim)
        new = q1d.copy()
        new[1:-1] -= F[1:] - F[:-1]
        cells = list(range(1, n - 1))
        if is_w:
            new[0] -= F[0]            # bottom: no flux through the ground
            new[-1] += F[-1]          # top: only the face below is antidiffusive
            cells = list(range(n))
        for c in cells:
            if c == 0:
                win = [0, 1]
                lo = min(q1d[0], q1d[1], q0d[0], q0d[1])
                hi = max(q1d[0], q1d[1], q0d[0], q0d[1])
            elif c == n - 1:
                lo = min(q1d[n - 2], q1d[n - 1], q0d[n - 2])
                hi = max(q1d[n - 2], q1d[n - 1], q0d[n - 2])
            else:
                win = [c - 1, c, c + 1]
                lo = min(q1d[win].min(), q0d[win].min())
                hi = max(q1d[win].max(), q0d[win].max())
            tol = 1e-6 * max(1.0, abs(hi))
            assert lo - tol <= new[c] <= hi + tol, (trial, c, lo, new[c], hi)


def test_fct_leaves_small_fluxes_alone(orc):
    """Where even the worst-case one-sided fluxes keep both cells inside their
    bounds, the limiter factor is 1."""
    n = 12
    q = 1.0 + 0.01 * np.arange(n)                 # smooth ramp
    q1 = q.copy()
    q0 = q + 0.5 * np.sin(np.arange(n))           # wide bounds
    U2 = np.full(n - 1, 1e-3)
    lim = orc.fct_core(q1, q0, U2, False)
    np.testing.assert_array_equal(lim[1:-1], orc.arr(U2)[1:-1])


def _boundary_case(end):
    """A 6-cell row where only the end cell's own bound could limit the end
    face.  The end cell sits 0.1 above its local minimum and sends 0.5 through
    the end face; the interior cell next to it has 0.6 of room for the 0.5 it
    receives (beta 1.2).  ``end`` 0 is the row start (flow +), 1 the row end
    (flow -, the mirror image)."""
    q1 = np.array([1.0, 0.9, 1.5, 1.5, 1.5, 1.5])
    U2 = np.array([0.5, 0.0, 0.0, 0.0, 0.0])
    if end == 1:
        q1, U2 = q1[::-1].copy(), -U2[::-1]
    return q1, U2


@pytest.mark.parametrize("end", [0, 1])
def test_fct_boundary_cell_is_not_limited_in_x_and_y(orc, end):
    """adv_mpdata_FCT_core.f90:79-83 and :95-99: in x and y the boundary cell
    has fin = fout = 0 ("no flux limitations to the boundary cell"), so its
    beta is (q1 - qmin) / 1e-15, about 1e14, and the end face keeps its full
    antidiffusive velocity even though the 0.5 it carries out of the end cell
    is five times that cell's 0.1 of room above its minimum."""
    q1, U2 = _boundary_case(end)
    lim = orc.fct_core(q1, q1, U2, False)
    face = 0 if end == 0 else -1
    assert lim[face] == orc.arr(U2)[face]
    np.testing.assert_array_equal(lim, orc.arr(U2))


@pytest.mark.parametrize("end", [0, 1])
def test_fct_boundary_cell_is_limited_in_w(orc, end):
    """adv_mpdata_FCT_core.f90:76-78 and :92-94: a w column's end cell counts
    its one face (fin = -min(0,f), fout = max(0,f)), so the same row limits
    the end face to the end cell's room: factor (q1 - qmin)/(0.5 + 1e-15),
    about 0.2, from the end cell (the interior cell's beta is 1.2)."""
    q1, U2 = _boundary_case(end)
    lim = orc.fct_core(q1, q1, U2, True).astype(np.float64)
    q = orc.arr(q1).astype(np.float64)
    face = 0 if end == 0 else -1
    cell, other = (0, 1) if end == 0 else (-1, -2)
    room = q[cell] - q[other]                       # q1 - qmin of the end cell
    assert 0.09 < room < 0.11
    exp = room / (0.5 + 1e-15) * orc.arr(U2).astype(np.float64)[face]
    np.testing.assert_allclose(lim[face], exp, rtol=1e-6)
    np.testing.assert_array_equal(lim[1:-1] if end == 0 else lim[:-1],
                                  orc.arr(U2)[1:-1] if end == 0 else orc.arr(U2)[:-1])


@pytest.mark.parametrize("end", [0, 1])
def test_fct_boundary_cell_at_its_minimum_stops_its_outflow(orc, end):
    """The other side of fin = fout = 0 for x and y: a boundary cell whose q1
    is its own local minimum has beta_out = 0 / 1e-15 = 0, so the face out of
    it is zeroed (factor min(1, beta_in, 0) = 0), however much room the cell
    beyond it has.  A restatement that skipped the boundary cell's beta
    instead would keep the face."""
    q1 = np.array([0.9, 1.0, 1.5, 1.5, 1.5, 1.5])
    U2 = np.array([0.5, 0.0, 0.0, 0.0, 0.0])
    if end == 1:
        q1, U2 = q1[::-1].copy(), -U2[::-1]
    lim = orc.fct_core(q1, q1, U2, False)
    face = 0 if end == 0 else -1
    assert lim[face] == 0.0


def run_test_mpdata(orc, fct, loops=10):
    case = synthetic.test_mpdata_case(fct=fct)
    d = case.domain
    q = {"qtr": orc.arr(d.species["qtr"])}
    stats = []
    for _ in range(loops):
        q = orc.run(d, case.options, case.dt, case.steps, case.bc_x, case.bc_y, q)
        a = q["qtr"][:, 1, 1].astype(np.float64)  # q(2,1,:) of the reference
        stats.append((a.mean(), a.max(), a.min()))
    return stats


def test_reference_step_function_scenario(orc):
    """test_mpdata.f90:51-111: a 1-to-2 step advected around a periodic 98-row
    channel at Courant 0.25, 10 loops of 392 calls.  The reference prints the
    mean, max and min and asserts nothing; with FCT the 1-D limiter bound
    keeps the field inside [1, 2], and the mean (1.5 at the start) drifts
    only through the one-cell periodic wrap of the second pass."""
    s = run_test_mpdata(orc, True)
    for mean, mx, mn in s:
        assert mx <= 2.0 + 1e-6 and mn >= 1.0 - 1e-6
        assert abs(mean - 1.5) < 1e-3
    # the step spreads: FCT-limited MPDATA is diffusive but not donor-cell
    assert 1.7 < s[-1][1] < 1.9


# --------------------------------------------------------------------------
# mpdata_advect3d (adv_mpdata.f90:352-413)
# --------------------------------------------------------------------------
def random_case(orc, nx, ny, nz, seed):
    r = rnd(seed)
    shape = (ny, nz, nx)
    q = orc.arr((r.random(shape) + 0.05) * (r.random(shape) > 0.3))
    U = orc.arr(r.uniform(-0.3, 0.3, shape))
    U[:, :, 0] = 0
    V = orc.arr(r.uniform(-0.3, 0.3, shape))
    V[0] = 0
    dzl = 50 * 1.15 ** np.arange(nz)
    dz = orc.arr(np.broadcast_to(dzl[None, :, None], shape))
    W = orc.arr(r.uniform(-0.3, 0.3, shape) * dz)
    rho = orc.arr(0.8 + 0.4 * r.random(shape))
    jaco = orc.arr(0.7 + 0.3 * r.random(shape))
    return q, U, V, W, rho, jaco, dz


def test_order1_equals_upwind(orc):
    q, U, V, W, rho, jaco, dz = random_case(orc, 10, 9, 6, 10)
    a = orc.advect3d(q, rho, jaco, dz, U, V, W, 1, True)
    b = orc.upwind(q, U, V, W, rho, dz, jaco)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(orc.advect3d(q, rho, jaco, dz, U, V, W, 0, True), q)


def _donor_1d(q, c, top=None):
    """Textbook donor cell along the last axis with Courant numbers c at the
    faces (c[..., m] is the face below cell m, m >= 1; for the vertical,
    face m is the top of cell m-1).  Without ``top`` the two end cells are
    not updated (the lateral ring); with ``top`` the first cell has a closed
    bottom and the last cell loses q_last * top through the lid."""
    F = np.zeros(q.shape[:-1] + (q.shape[-1] + 1,))
    F[..., 1:-1] = np.maximum(c[..., 1:], 0) * q[..., :-1] + np.minimum(c[..., 1:], 0) * q[..., 1:]
    out = q.copy()
    if top is None:
        out[..., 1:-1] = q[..., 1:-1] - (F[..., 2:-1] - F[..., 1:-2])
    else:
        F[..., -1] = q[..., -1] * top
        out = q - (F[..., 1:] - F[..., :-1])
    return out


def _antidiff_1d(q, c):
    """Smolarkiewicz (1983) antidiffusive Courant number at every interior
    face, (|c| - c^2) (r - l) / (r + l + 1e-10)."""
    a = np.zeros_like(c)
    r, l = q[..., 1:], q[..., :-1]
    a[..., 1:] = (np.abs(c[..., 1:]) - c[..., 1:] ** 2) * (r - l) / (r + l + 1e-10)
    return a


@pytest.mark.parametrize("axis", ["x", "z"])
def test_order2_1d_is_textbook_mpdata_halved(orc, axis):
    """G = 1, one wind component: order 2 without FCT is two textbook
    donor-cell passes, the second with the antidiffusive Courant number times
    0.5 (:378-380).  In z the reference's W is a Courant number times dz and
    the pseudo-velocity is formed from W/dz and scaled back by dz (:374,380),
    so a uniform dz = 2.5 must drop out; the lid flux of pass 1 is
    q(top)*W(top) and that of pass 2 is zero (:96-97, :211-212).  The
    expectation is an independent FP64 NumPy restatement."""
    nx, ny, nz = 12, 4, 9
    shape = (ny, nz, nx)
    r = rnd(21)
    q = orc.arr(r.random(shape) + 0.2)
    U, V, W = (zeros(orc, shape) for _ in range(3))
    dzv = 2.5
    G, dz, rho = fields(orc, shape, dz=dzv)
    q64 = q.astype(np.float64)
    if axis == "x":
        U[:, :, 1:] = r.uniform(-0.8, 0.8, (ny, nz, nx - 1))
        c = U.astype(np.float64)
        q1 = _donor_1d(q64, c)
        q1[0], q1[-1] = q64[0], q64[-1]  # rows j = jms, jme are the ring
        exp = _donor_1d(q1, 0.5 * _antidiff_1d(q1, c))
        exp[0], exp[-1] = q64[0], q64[-1]
    else:
        W[...] = r.uniform(-0.8, 0.8, shape) * dzv
        cw = W.astype(np.float64) / dzv  # Courant number of the top face of each level
        qt = np.transpose(q64, (0, 2, 1))  # (ny, nx, nz): levels last
        c = np.zeros(qt.shape)
        c[..., 1:] = np.transpose(cw, (0, 2, 1))[..., :-1]
        q1 = _donor_1d(qt, c, top=np.transpose(cw, (0, 2, 1))[..., -1])
        exp = _donor_1d(q1, 0.5 * _antidiff_1d(q1, c), top=np.zeros(qt.shape[:-1]))
        exp = np.transpose(exp, (0, 2, 1))
        exp[0], exp[-1], exp[:, :, 0], exp[:, :, -1] = q64[0], q64[-1], q64[:, :, 0], q64[:, :, -1]
    out = orc.advect3d(q, rho, G, dz, U, V, W, 2, False).astype(np.float64)
    tol = 1e-12 if orc.dtype == np.float64 else 3e-6
    np.testing.assert_allclose(out, exp, rtol=0, atol=tol * np.abs(exp).max())
    # without the 0.5 damping the answer moves well beyond the tolerance
    if axis == "x":
        undamped = _donor_1d(q1, _antidiff_1d(q1, c))
        undamped[0], undamped[-1] = q64[0], q64[-1]
        assert np.abs(undamped - exp).max() > 1e3 * tol * np.abs(exp).max()


@pytest.mark.parametrize("fct", [True, False])
def test_order3_is_order2_then_one_more_corrective_pass(orc, fct):
    """The iord >= 3 quirk (:374,382,388-397): the third pass forms its
    pseudo-velocities from the physical Courant numbers, and its FCT bounds
    take q0 = q1 = the order-2 result.  Composed from the separately pinned
    functions; using the original field for q0 or the previous pass's
    pseudo-velocities fails it."""
    q, U, V, W, rho, jaco, dz = random_case(orc, 10, 8, 6, 22)
    r2 = orc.advect3d(q, rho, jaco, dz, U, V, W, 2, fct)
    u2, v2, w2 = scaled_fluxes(orc, r2, U, V, W, jaco * rho, dz)
    if fct:
        u2, v2, w2 = orc.flux_limiter(r2, r2, u2, v2, w2)
    exp = orc.upwind(r2, u2, v2, w2, rho, dz, jaco)
    np.testing.assert_array_equal(orc.advect3d(q, rho, jaco, dz, U, V, W, 3, fct), exp)
    if fct:  # q0 = the original field gives a different answer
        u2b, v2b, w2b = scaled_fluxes(orc, r2, U, V, W, jaco * rho, dz)
        u2b, v2b, w2b = orc.flux_limiter(q, r2, u2b, v2b, w2b)
        wrong = orc.upwind(r2, u2b, v2b, w2b, rho, dz, jaco)
        assert np.abs(wrong - exp).max() > 0


@pytest.mark.parametrize("order,fct", [(2, True), (3, True), (2, False)])
def test_x_and_y_mirror_symmetry(orc, order, fct):
    q, U, V, W, rho, jaco, dz = random_case(orc, 11, 9, 6, 11)
    base = orc.advect3d(q, rho, jaco, dz, U, V, W, order, fct)
    # mirror x: cell i -> nx-1-i, face i -> nx-i with U negated
    f = lambda a: np.ascontiguousarray(a[:, :, ::-1])
    Um = np.zeros_like(U)
    Um[:, :, 1:] = -U[:, :, 1:][:, :, ::-1]
    mx = orc.advect3d(f(q), f(rho), f(jaco), f(dz), Um, f(V), f(W), order, fct)
    scale = np.abs(base).max()
    np.testing.assert_allclose(f(mx), base, rtol=0, atol=2e-6 * scale)
    # mirror y
    g = lambda a: np.ascontiguousarray(a[::-1])
    Vm = np.zeros_like(V)
    Vm[1:] = -V[1:][::-1]
    my = orc.advect3d(g(q), g(rho), g(jaco), g(dz), g(U), Vm, g(W), order, fct)
    np.testing.assert_allclose(g(my), base, rtol=0, atol=2e-6 * scale)


@pytest.mark.parametrize("order", [2, 3])
def test_xy_transposition_symmetry(orc, order):
    """Swapping the x and y axes (and U with V) transposes the result on the
    cells advect updates: the x and y code paths are the same operator."""
    q, U, V, W, rho, jaco, dz = random_case(orc, 10, 10, 5, 12)
    base = orc.advect3d(q, rho, jaco, dz, U, V, W, order, True)
    t = lambda a: np.ascontiguousarray(np.transpose(a, (2, 1, 0)))
    tr = orc.advect3d(t(q), t(rho), t(jaco), t(dz), t(V), t(U), t(W), order, True)
    scale = np.abs(base).max()
    inner = (slice(1, -1), slice(None), slice(1, -1))
    np.testing.assert_allclose(t(tr)[inner], base[inner], rtol=0, atol=2e-6 * scale)


# --------------------------------------------------------------------------
# advect (advection_driver.f90:59-126)
# --------------------------------------------------------------------------
def tiny_domain(seed=13):
    case = synthetic.config1(nx=16, ny=7, nz=6, steps=1)
    r = rnd(seed)
    d = case.domain
    case.options.vars_to_advect = ["qv", "qtr"]
    d.set_var_lists(case.options, {"qv": d.species["qv"], "qtr": r.random(d.shape)})
    return case


def test_advect_rk3_is_a_noop(orc):
    case = tiny_domain()
    case.options.time_options.RK3 = True
    out = orc.advect(case.domain, case.options, case.dt)
    for n, a in out.items():
        np.testing.assert_array_equal(a, orc.arr(case.domain.species[n]))


def test_advect_species_are_independent(orc):
    case = tiny_domain()
    d = case.domain
    both = orc.advect(d, case.options, case.dt)
    one = orc.advect(d, case.options, case.dt, {"qtr": d.species["qtr"]})
    np.testing.assert_array_equal(both["qtr"], one["qtr"])


def test_density_shadowing(orc):
    """advect_density=.false.: the Courant numbers use rho = 1 (:489) but the
    divergence divides by the domain density passed as the dummy rho
    (advection_driver.f90:120).  u = 1, dt = dx: Courant 1 with density 2
    gives the average of a cell and its upwind neighbour, not a shift."""
    nx, ny, nz = 8, 5, 4
    d = Domain(nx, ny, nz, 10.0)
    d.u[...] = 1.0
    d.density[...] = 2.0
    opts = Options()
    opts.adv_options.mpdata_order = 1
    q = rnd(14).random(d.shape) + 1
    out = orc.advect(d, opts, 10.0, {"q": q})["q"]
    q = orc.arr(q)
    exp = q - (q - np.roll(q, 1, axis=2)) / 2
    inner = (slice(1, -1), slice(None), slice(1, -1))
    np.testing.assert_allclose(out[inner], exp[inner], rtol=1e-6)
    opts.parameters.advect_density = True
    out2 = orc.advect(d, opts, 10.0, {"q": q})["q"]
    np.testing.assert_allclose(out2[inner], np.roll(q, 1, axis=2)[inner], rtol=1e-6)


def test_divergence_and_mass_closed_forms(orc):
    nx, ny, nz = 6, 5, 4
    shape = (ny, nz, nx)
    U = zeros(orc, shape)
    V = zeros(orc, shape)
    W = zeros(orc, shape)
    U[:, :, 3] = 0.01  # face between cells 2 and 3: du = +0.01 at i=2, -0.01 at i=3
    n, mx = orc.divergence(U, V, W, orc.arr(np.ones(shape)))
    assert n == 2 * (ny - 2) * nz and abs(mx - 0.01) < 1e-6
    q = orc.arr(np.full(shape, 2.0))
    m = orc.mass(q, orc.arr(np.full(shape, 0.5)), orc.arr(np.full(shape, 3.0)),
                 orc.arr(np.full(shape, 4.0)))
    assert m == pytest.approx(2.0 * 0.5 * 3.0 * 4.0 * (nx - 2) * nz * (ny - 2))


def test_apply_bc(orc):
    q = orc.arr(rnd(15).random((6, 3, 7)))
    b = orc.apply_bc(q.copy(), 1, 2)
    np.testing.assert_array_equal(b[1:-1, :, -1], q[1:-1, :, -2])
    np.testing.assert_array_equal(b[0], b[-2])
    np.testing.assert_array_equal(b[-1], b[1])
