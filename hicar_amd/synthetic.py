"""Synthetic domains for the MPDATA benchmarks and parity cases.

Every random draw comes from a counter-based generator (Philox) keyed by
(config, domain, field), so any machine with this numpy produces the same
inputs.  Geometry follows ICAR: terrain-following jacobian (H-h)/H on every
level (domain_obj.f90:1787-1799 with max_level = nz), staggered jacobians as
domain_obj.f90:1916-1931, advection_dz = dz_levels(k) (:1960-1962), and dt
from the cfl_strictness=3 measure max(|u|/dx, |v|/dx, |w|/dz) (time_step.f90:
98-127,156).  The vertical wind is diagnosed from the discrete continuity of
the MPDATA Courant numbers themselves, so the flow is non-divergent for the
scheme (a constant field stays constant to rounding).
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .domain import Domain, Options, kADV_STD, kFLUXCOR_MONO, halo_size

HYDRO_MASS = ("qc", "qr", "qs", "qi", "qg")
NUMBERS = ("ni", "nr", "ns", "ng")


def rng(config: str, domain: int, name: str) -> np.random.Generator:
    key = hashlib.sha256(f"{config}/{domain}/{name}".encode()).digest()
    return np.random.Generator(np.random.Philox(int.from_bytes(key[:8], "little")))


def dz_levels(nz: int) -> np.ndarray:
    """dz_k = 50 m * 1.15^k."""
    return 50.0 * 1.15 ** np.arange(nz, dtype=np.float64)


def fractal_terrain(nx: int, ny: int, lo: float, hi: float, r: np.random.Generator,
                    slope: float = -2.0) -> np.ndarray:
    """Periodic random surface with power spectrum ~ |k|^slope, rescaled to
    [lo, hi]; shape (ny, nx), float64."""
    noise = r.standard_normal((ny, nx))
    kx = np.fft.fftfreq(nx)[None, :]
    ky = np.fft.fftfreq(ny)[:, None]
    kk = np.sqrt(kx * kx + ky * ky)
    kk[0, 0] = np.inf
    spec = np.fft.fft2(noise) * kk ** (slope / 2.0)
    h = np.real(np.fft.ifft2(spec))
    h -= h.min()
    h *= (hi - lo) / max(h.max(), 1e-30)
    return h + lo


@dataclass
class Case:
    """A synthetic workload: the domain, its options and how to step it."""
    name: str
    domain: Domain
    options: Options
    dt: float
    steps: int
    bc_x: int = 1
    bc_y: int = 2
    meta: dict = field(default_factory=dict)


def _cap(h: float, nz: int) -> float:
    """Terrain height, capped at 40 % of the model top for the small grids
    of the tests (the full configurations are far below the cap)."""
    return min(h, 0.4 * float(dz_levels(nz).sum()))


def _metric(dom: Domain, terrain: np.ndarray, dz: np.ndarray) -> np.ndarray:
    """Fills jacobians and advection_dz; returns cell-centre heights z."""
    ny, nz, nx = dom.shape
    H = dz.sum()
    jac2 = (H - terrain) / H                                   # (ny, nx)
    jac = np.broadcast_to(jac2[:, None, :], (ny, nz, nx))
    dom.jacobian[...] = jac
    ju = np.empty((ny, nz, nx + 1))
    ju[:, :, 1:nx] = 0.5 * (jac[:, :, 1:] + jac[:, :, :-1])
    ju[:, :, 0] = jac[:, :, 0]
    ju[:, :, nx] = jac[:, :, nx - 1]
    dom.jacobian_u[...] = ju
    jv = np.empty((ny + 1, nz, nx))
    jv[1:ny] = 0.5 * (jac[1:] + jac[:-1])
    jv[0] = jac[0]
    jv[ny] = jac[ny - 1]
    dom.jacobian_v[...] = jv
    dzi = dz[None, :, None] * jac
    jw = np.empty((ny, nz, nx))
    jw[:, :nz - 1] = (dzi[:, :-1] * jac[:, :-1] + dzi[:, 1:] * jac[:, 1:]) / (dzi[:, :-1] + dzi[:, 1:])
    jw[:, nz - 1] = jac[:, nz - 1]
    dom.jacobian_w[...] = jw
    dom.advection_dz[...] = dz[None, :, None]
    zc = np.cumsum(dz) - 0.5 * dz
    return terrain[:, None, :] + jac * zc[None, :, None]


def _face_x(a: np.ndarray) -> np.ndarray:
    """Cell values (ny,nz,nx) -> x-face values (ny,nz,nx+1), ends copied."""
    f = np.empty(a.shape[:-1] + (a.shape[-1] + 1,))
    f[..., 1:-1] = 0.5 * (a[..., 1:] + a[..., :-1])
    f[..., 0] = a[..., 0]
    f[..., -1] = a[..., -1]
    return f


def _face_y(a: np.ndarray) -> np.ndarray:
    f = np.empty((a.shape[0] + 1,) + a.shape[1:])
    f[1:-1] = 0.5 * (a[1:] + a[:-1])
    f[0] = a[0]
    f[-1] = a[-1]
    return f


def _diagnose_w(dom: Domain, u: np.ndarray, v: np.ndarray, advect_density: bool,
                std: bool = False) -> np.ndarray:
    """w such that the MPDATA Courant numbers of mpdata_compute_wind
    (adv_mpdata.f90:492-498) satisfy
        (U(i+1)-U(i)) + (V(j+1)-V(j)) + (W(k)-W(k-1))/dz_k = 0,  W(-1) = 0,
    in float64 (per unit dt).  With ``std`` the same for the mass fluxes of
    adv_std_compute_wind (advect.f90:475-484), whose interface density is
    dz-weighted, as balance_uvw gives it (wind.f90:177)."""
    ny, nz, nx = dom.shape
    rho = dom.density.astype(np.float64) if advect_density else np.ones((ny, nz, nx))
    ru = _face_x(rho)
    rv = _face_y(rho)
    Ur = u * ru * dom.jacobian_u / dom.dx
    Vr = v * rv * dom.jacobian_v / dom.dx
    divh = (Ur[:, :, 1:] - Ur[:, :, :-1]) + (Vr[1:] - Vr[:-1])
    dz = dom.advection_dz.astype(np.float64)
    Wr = -np.cumsum(dz * divh, axis=1)
    rz = np.empty_like(rho)
    if std:
        rz[:, :nz - 1] = (rho[:, :-1] * dz[:, :-1] + rho[:, 1:] * dz[:, 1:]) / (dz[:, :-1] + dz[:, 1:])
    else:
        rz[:, :nz - 1] = 0.5 * (rho[:, 1:] + rho[:, :-1])
    rz[:, nz - 1] = rho[:, nz - 1]
    return Wr / (dom.jacobian_w * rz)


def courant_dt(dom: Domain, u: np.ndarray, v: np.ndarray, w: np.ndarray,
               cfl: float = 0.5) -> float:
    """dt = cfl / max(|u|/dx, |v|/dx, |w|/dz) in float64, returned as the
    float32 value advect receives."""
    dz = dom.advection_dz.astype(np.float64)
    mx = max(np.abs(u).max() / dom.dx, np.abs(v).max() / dom.dx, (np.abs(w) / dz).max())
    return float(np.float32(cfl / mx))


def _lognormal(r, shape, mean, sigma=1.5):
    mu = np.log(mean) - 0.5 * sigma * sigma
    return np.exp(mu + sigma * r.standard_normal(shape))


def _species(config: str, d: int, name: str, shape, z: np.ndarray,
             zero_frac: float = 0.6, qv_profile: bool = False) -> np.ndarray:
    r = rng(config, d, name)
    if name == "theta":
        return 290.0 + 4e-3 * z + 0.2 * r.standard_normal(shape)
    if name == "qv" and qv_profile:
        return 0.012 * np.exp(-z / 2500.0) * (1.0 + 0.05 * r.standard_normal(shape))
    if name in NUMBERS:
        return _lognormal(r, shape, 1e5)
    q = _lognormal(r, shape, 1e-4)
    mask = rng(config, d, name + "/mask").random(shape) < zero_frac
    q[mask] = 0.0
    return q


def _finish(name: str, dom: Domain, opts: Options, u, v, density, names, z, steps,
            bc_x, bc_y, cfg_key, d_index=0, qv_profile=False, meta=None, cfl=0.5) -> Case:
    dom.density[...] = density
    w = _diagnose_w(dom, u, v, opts.parameters.advect_density,
                    std=opts.physics.advection == kADV_STD)
    dom.u[...] = u
    dom.v[...] = v
    dom.w[...] = w
    dt = courant_dt(dom, u, v, w, cfl)
    opts.vars_to_advect = list(names)
    values = {n: _species(cfg_key, d_index, n, dom.shape, z, qv_profile=qv_profile)
              for n in names}
    dom.set_var_lists(opts, values)
    return Case(name, dom, opts, dt, steps, bc_x, bc_y, meta or {})


def _ridge_winds(dom: Domain, z: np.ndarray):
    """u = 10 m/s + 2 m/s sin(2 pi x / 64 km) exp(-z / 3 km), v = 0."""
    ny, nz, nx = dom.shape
    xf = (np.arange(nx + 1) - 0.5) * dom.dx
    zf = _face_x(z)
    u = 10.0 + 2.0 * np.sin(2 * np.pi * xf / 64e3)[None, None, :] * np.exp(-zf / 3000.0)
    v = np.zeros((ny + 1, nz, nx))
    return u, v


def config1(nx=256, ny=64, nz=20, steps=100) -> Case:
    """Plain 2-pass MPDATA, FP32, qv only, no FCT; Gaussian ridge (1500 m,
    half-width 10 km) at mid-x; periodic y, fixed inflow, zero-gradient
    outflow."""
    dom = Domain(nx, ny, nz, 1000.0)
    x = np.arange(nx) * dom.dx
    xm = 0.5 * (nx - 1) * dom.dx
    terrain = np.broadcast_to(_cap(1500.0, nz) * np.exp(-((x - xm) / 10e3) ** 2), (ny, nx))
    z = _metric(dom, terrain, dz_levels(nz))
    opts = Options()
    opts.adv_options.mpdata_order = 2
    opts.adv_options.MPDATA_FCT = False
    opts.parameters.advect_density = True
    u, v = _ridge_winds(dom, z)
    return _finish("C1", dom, opts, u, v, 1.2 * np.exp(-z / 8000.0), ["qv"], z, steps, 1, 2,
                   "C1", qv_profile=True)


def _fractal_winds(dom: Domain, terrain: np.ndarray, z: np.ndarray, U0: float):
    """U0 westerly plus a terrain-forced perturbation 0.2*U0*grad(h),
    decaying over 3 km above the ground."""
    ny, nz, nx = dom.shape
    dhdx = (np.roll(terrain, -1, axis=1) - np.roll(terrain, 1, axis=1)) / (2 * dom.dx)
    dhdy = (np.roll(terrain, -1, axis=0) - np.roll(terrain, 1, axis=0)) / (2 * dom.dx)
    zagl = z - terrain[:, None, :]
    decay = np.exp(-zagl / 3000.0)
    u = _face_x(U0 + 0.2 * U0 * dhdx[:, None, :] * decay)
    v = _face_y(0.2 * U0 * dhdy[:, None, :] * decay)
    return u, v


C2_SPECIES = ("qv", "theta", "qc", "qr", "qs", "qi", "qg", "ni", "nr", "qtr")


def config2(nx=1024, ny=1024, nz=30, steps=50, nspecies=10) -> Case:
    """Single realistic domain: 3-pass MPDATA with FCT, 10 species, fractal
    terrain 0-3000 m, 15 m/s westerly with terrain-forced perturbation."""
    dom = Domain(nx, ny, nz, 1000.0)
    terrain = fractal_terrain(nx, ny, 0.0, _cap(3000.0, nz), rng("C2", 0, "terrain"))
    z = _metric(dom, terrain, dz_levels(nz))
    opts = Options()
    opts.adv_options.mpdata_order = 3
    opts.adv_options.MPDATA_FCT = True
    opts.parameters.advect_density = True
    u, v = _fractal_winds(dom, terrain, z, 15.0)
    return _finish("C2", dom, opts, u, v, 1.2 * np.exp(-z / 8000.0), C2_SPECIES[:nspecies], z,
                   steps, 1, 2, "C2")


def config_std(nx=1024, ny=1024, nz=30, steps=50, nspecies=10, h_order=5, v_order=3,
               flux_corr=1, rk3=True, cfl=1.2) -> Case:
    """The C2 domain (fractal terrain, 15 m/s westerly with terrain-forced
    perturbation, the same species) under the standard scheme with the
    settings of run/complete_icar_options.nml: h_order 5, v_order 3, the WRF
    limiter and RK3 at a Courant number of 1.2 (below the 1.3 of that
    namelist, which options_check caps at 1.4 for order 5).  The halo is
    the one ICAR gives these options; w balances the scheme's mass fluxes."""
    opts = Options()
    opts.physics.advection = kADV_STD
    opts.adv_options.h_order = h_order
    opts.adv_options.v_order = v_order
    opts.adv_options.flux_corr = flux_corr
    opts.time_options.RK3 = rk3
    opts.parameters.advect_density = True
    dom = Domain(nx, ny, nz, 1000.0, halo=halo_size(opts))
    terrain = fractal_terrain(nx, ny, 0.0, _cap(3000.0, nz), rng("C2", 0, "terrain"))
    z = _metric(dom, terrain, dz_levels(nz))
    u, v = _fractal_winds(dom, terrain, z, 15.0)
    return _finish("STD", dom, opts, u, v, 1.2 * np.exp(-z / 8000.0), C2_SPECIES[:nspecies], z,
                   steps, 1, 2, "C2", cfl=cfl)


C3_SPECIES = ("qv", "qc", "qr", "qs", "qi", "qg")


def config3_domain(d: int, steps=100, lo: int = 64, hi: int = 512, nzs=(20, 30, 40)) -> Case:
    """Domain d of the 48-member ensemble: nx, ny ~ U[lo,hi] (64..512), nz
    from nzs ({20,30,40}), its own fractal terrain, C1-type winds, 6 FP32
    species, order 2 with FCT (namelist defaults)."""
    r = rng("C3", d, "shape")
    nx = int(r.integers(lo, hi + 1))
    ny = int(r.integers(lo, hi + 1))
    nz = int(nzs[int(r.integers(0, len(nzs)))])
    dom = Domain(nx, ny, nz, 1000.0)
    terrain = fractal_terrain(nx, ny, 0.0, _cap(3000.0, nz), rng("C3", d, "terrain"))
    z = _metric(dom, terrain, dz_levels(nz))
    opts = Options()
    opts.parameters.advect_density = True
    opts.kernel_variant = 1
    u, v = _ridge_winds(dom, z)
    return _finish(f"C3[{d}]", dom, opts, u, v, 1.2 * np.exp(-z / 8000.0), C3_SPECIES, z, steps,
                   1, 2, "C3", d_index=d)


def config3(n: int = 48, steps=100, lo: int = 64, hi: int = 512, nzs=(20, 30, 40)) -> list:
    """The ensemble: n independent domains of uneven sizes."""
    return [config3_domain(d, steps, lo, hi, nzs) for d in range(n)]


C4_SPECIES = ("qv", "theta", "qc", "qr", "qs", "qi", "qg", "ni", "nr", "ns", "ng", "qtr")


def pow2_scale(a: np.ndarray, storage: int) -> int:
    """Per-species power-of-two exponent e (device holds q*2^-e): the
    largest magnitude maps just below 2^15 for FP16, to [0.5,1) for BF16."""
    m = float(np.abs(a).max())
    if m == 0.0:
        return 0
    e = int(np.ceil(np.log2(m)))
    return e - 15 if storage in (_lib.STORE_F16, _lib.STORE_F16_SR) else e


def config4(storage: int = _lib.STORE_BF16, nx=256, ny=64, nz=20, steps=200) -> Case:
    """The C1 domain with 12 species; hydrometeor masses and numbers stored
    in BF16 or FP16 with per-species power-of-two scaling, qv and theta in
    FP32; order 2 with FCT."""
    dom = Domain(nx, ny, nz, 1000.0)
    x = np.arange(nx) * dom.dx
    xm = 0.5 * (nx - 1) * dom.dx
    terrain = np.broadcast_to(_cap(1500.0, nz) * np.exp(-((x - xm) / 10e3) ** 2), (ny, nx))
    z = _metric(dom, terrain, dz_levels(nz))
    opts = Options()
    opts.parameters.advect_density = True
    opts.kernel_variant = 1
    u, v = _ridge_winds(dom, z)
    case = _finish("C4", dom, opts, u, v, 1.2 * np.exp(-z / 8000.0), C4_SPECIES, z, steps, 1, 2,
                   "C4", qv_profile=True)
    for n, a in dom.species.items():
        if n not in ("qv", "theta"):
            dom.storage[n] = (storage, pow2_scale(a, storage))
    return case


def dz_levels5(nz: int) -> np.ndarray:
    """dz_k = 50 m * 1.07^k: 50 levels reach 20.3 km (the C1 stretching,
    1.15^k, would put the 50th level at 384 km)."""
    return 50.0 * 1.07 ** np.arange(nz, dtype=np.float64)


def config5(nx=4096, ny=4096, nz=50, steps=20, storage: int = _lib.STORE_BF16_SR,
            workers: int = 16, log=None, scheme: str = "mpdata", names=None) -> Case:
    """Large single domain, batched and mixed precision: dx = 250 m, fractal
    terrain 0-4000 m with a 3000 m ridge line (Gaussian, 5 km half-width, at
    a third of x), a 20 m/s westerly jet at 8 km (3 km wide) plus a
    terrain-forced perturbation 0.2*U0*grad(h) of the terrain smoothed over
    2 km, both in mass-flux form (u*rho*J is the level's mean rho*J times the
    wind) with a perturbation profile that returns aloft what it carries
    near the ground, w from continuity, Courant 0.5;
    the 12 species of config 4, hydrometeor masses and numbers and the tracer
    stored at 16 bits with per-species power-of-two scaling (BF16 with
    stochastic rounding by default: unbiased, so the mass stays within its
    rounding's random walk); order 2 with FCT.  At 4096 x 4096 x 50 (839 M cells) the arrays are filled one level
    at a time, levels in parallel threads, so that no float64 3-D temporary
    exists; the formulas are those of the other configurations, level by
    level (the random streams are keyed by level).  With scheme "std" the
    same domain is set up for the standard scheme (h_order 5, v_order 3, the
    WRF limiter, RK3; all species FP32, `names` of them), w balancing its
    mass fluxes with their dz-weighted interface density."""
    from concurrent.futures import ThreadPoolExecutor

    say = log or (lambda m: None)
    std = scheme == "std"
    opts = Options()
    opts.parameters.advect_density = True
    if std:
        opts.physics.advection = kADV_STD
        opts.adv_options.h_order = 5
        opts.adv_options.v_order = 3
        opts.adv_options.flux_corr = kFLUXCOR_MONO
        opts.time_options.RK3 = True
    else:
        opts.adv_options.mpdata_order = 2
        opts.adv_options.MPDATA_FCT = True
        opts.kernel_variant = 1
    names = list(names or C4_SPECIES)
    opts.vars_to_advect = names
    dom = Domain(nx, ny, nz, 250.0, halo=halo_size(opts))
    dz = dz_levels5(nz)
    H = dz.sum()
    zc = np.cumsum(dz) - 0.5 * dz
    x = np.arange(nx) * dom.dx
    ridge = 3000.0 * np.exp(-((x - nx * dom.dx / 3.0) / 5e3) ** 2)
    terrain = np.maximum(fractal_terrain(nx, ny, 0.0, 4000.0, rng("C5", 0, "terrain")),
                         ridge[None, :])
    jac = (H - terrain) / H                                   # (ny, nx), every level
    ju = np.empty((ny, nx + 1))
    ju[:, 1:nx] = 0.5 * (jac[:, 1:] + jac[:, :-1])
    ju[:, 0] = jac[:, 0]
    ju[:, nx] = jac[:, nx - 1]
    jv = np.empty((ny + 1, nx))
    jv[1:ny] = 0.5 * (jac[1:] + jac[:-1])
    jv[0] = jac[0]
    jv[ny] = jac[ny - 1]
    # the wind perturbation follows the terrain smoothed over 2 km
    kx = np.fft.fftfreq(nx)[None, :]
    ky = np.fft.fftfreq(ny)[:, None]
    sig = 2000.0 / dom.dx
    hs = np.real(np.fft.ifft2(np.fft.fft2(terrain) *
                              np.exp(-2.0 * (np.pi * sig) ** 2 * (kx * kx + ky * ky))))
    dhdx = (np.roll(hs, -1, axis=1) - np.roll(hs, 1, axis=1)) / (2 * dom.dx)
    dhdy = (np.roll(hs, -1, axis=0) - np.roll(hs, 1, axis=0)) / (2 * dom.dx)
    U0 = 20.0

    def z_of(k):
        return terrain + jac * zc[k]

    # R_k: the level mean of rho*J; s_k: the perturbation's vertical profile,
    # exp(-z/3 km) less a constant such that sum_k dz_k R_k s_k = 0, so the
    # terrain-forced inflow below returns aloft and W closes at the top
    R = np.array([float(np.mean(1.2 * np.exp(-z_of(k) / 8000.0) * jac)) for k in range(nz)])
    e = np.exp(-zc / 3000.0)
    prof = e - float(np.sum(dz * R * e) / np.sum(dz * R))

    def level_fields(k):
        """Metric, density and horizontal winds of level k (as _metric,
        _fractal_winds with a jet profile); returns the level's horizontal
        mass-flux divergence and the interface density for w."""
        z = z_of(k)
        dom.jacobian[:, k, :] = jac
        dom.jacobian_u[:, k, :] = ju
        dom.jacobian_v[:, k, :] = jv
        dom.advection_dz[:, k, :] = dz[k]
        if k < nz - 1:
            dzi, dzu = dz[k] * jac, dz[k + 1] * jac
            dom.jacobian_w[:, k, :] = (dzi * jac + dzu * jac) / (dzi + dzu)
        else:
            dom.jacobian_w[:, k, :] = jac
        dom.density[:, k, :] = 1.2 * np.exp(-z / 8000.0)
        rho = dom.density[:, k, :].astype(np.float64)  # the device's density
        # mass-flux form: u*rho*J = R_k * (jet + 0.2*U0*dh/dx*s_k), the jet
        # 20 m/s at 8 km (3 km wide) where rho*J equals its level mean
        jet = U0 * np.exp(-((zc[k] - 8000.0) / 3000.0) ** 2)
        ru = _face_x(rho) * dom.jacobian_u[:, k, :]
        rv = _face_y(rho) * dom.jacobian_v[:, k, :]
        u = R[k] * (jet + _face_x(0.2 * U0 * dhdx * prof[k])) / ru
        v = R[k] * _face_y(0.2 * U0 * dhdy * prof[k]) / rv
        dom.u[:, k, :] = u
        dom.v[:, k, :] = v
        Ur = u * _face_x(rho) * dom.jacobian_u[:, k, :] / dom.dx
        Vr = v * _face_y(rho) * dom.jacobian_v[:, k, :] / dom.dx
        divh = (Ur[:, 1:] - Ur[:, :-1]) + (Vr[1:] - Vr[:-1])
        if k < nz - 1:
            ra = np.float32(1.2 * np.exp(-z_of(k + 1) / 8000.0)).astype(np.float64)
            if std:  # adv_std_compute_wind's interface density (advect.f90:475-476)
                dk, da = np.float64(np.float32(dz[k])), np.float64(np.float32(dz[k + 1]))
                rz = (rho * dk + ra * da) / (dk + da)
            else:
                rz = 0.5 * (ra + rho)
        else:
            rz = rho
        umax = max(np.abs(u).max(), np.abs(v).max()) / dom.dx
        return divh, rz, umax

    say(f"C5 {nx}x{ny}x{nz}: metric, density and winds ...")
    with ThreadPoolExecutor(workers) as ex:
        lev = list(ex.map(level_fields, range(nz)))
    # w from continuity, bottom up (as _diagnose_w)
    Wr = np.zeros((ny, nx))
    cmax = max(l[2] for l in lev)
    jw = dom.jacobian_w
    for k in range(nz):
        Wr = Wr - dz[k] * lev[k][0]
        w = Wr / (jw[:, k, :].astype(np.float64) * lev[k][1])
        dom.w[:, k, :] = w
        cmax = max(cmax, np.abs(w).max() / dz[k])
    del lev
    dt = float(np.float32(0.5 / cmax))

    def species_level(args):
        name, k = args
        r = rng("C5", k, name)
        shape = (ny, nx)
        z = z_of(k)
        if name == "theta":
            return 290.0 + 4e-3 * z + 0.2 * r.standard_normal(shape)
        if name == "qv":
            return 0.012 * np.exp(-z / 2500.0) * (1.0 + 0.05 * r.standard_normal(shape))
        if name in NUMBERS:
            return _lognormal(r, shape, 1e5)
        q = _lognormal(r, shape, 1e-4)
        q[rng("C5", k, name + "/mask").random(shape) < 0.6] = 0.0
        return q

    values = {}
    for name in names:
        say(f"C5 species {name} ...")
        a = np.empty(dom.shape, np.float32)

        def fill(k, name=name, a=a):
            a[:, k, :] = species_level((name, k))

        with ThreadPoolExecutor(workers) as ex:
            list(ex.map(fill, range(nz)))
        values[name] = a
    dom.set_var_lists(opts, values)
    del values
    if not std:
        for n, a in dom.species.items():
            if n not in ("qv", "theta"):
                dom.storage[n] = (storage, pow2_scale(a, storage))
    return Case("C5", dom, opts, dt, steps, 1, 2, {"dz": "50*1.07^k", "scheme": scheme})


def test_mpdata_case(nloops_steps: int = 392, fct: bool = True) -> Case:
    """The scenario of src/tests/test_mpdata.f90:51-111: nx=3, ny=100, nz=3,
    v = cfl = 0.25, u = w = 0, dz = rho = 1, mpdata_order 2, a step from 1
    to 2 at mid-y, periodic wrap of rows 1 and ny after every call."""
    nx, ny, nz = 3, 100, 3
    dom = Domain(nx, ny, nz, 1.0)
    dom.v[...] = 0.25
    opts = Options()
    opts.adv_options.mpdata_order = 2
    opts.adv_options.MPDATA_FCT = fct
    q = np.ones((ny, nz, nx))
    q[ny // 2:] = 2.0  # Fortran rows i > ny/2 (1-based)
    q[0] = q[ny - 2]
    q[ny - 1] = q[1]
    opts.vars_to_advect = ["qtr"]
    dom.set_var_lists(opts, {"qtr": q})
    return Case("test_mpdata", dom, opts, 1.0, nloops_steps, 0, 2)
