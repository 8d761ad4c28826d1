"""ctypes wrapper of the C oracle (oracle/mpdata_oracle.c, std_oracle.c,
dyn_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, never by the product package.  Arrays follow
hicar_amd.domain: C-order (ny, nz, nx) numpy arrays.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class Oracle:
    """FP32 (dtype=np.float32) or FP64 (np.float64) restatement."""

    def __init__(self, dtype=np.float32, fma: bool = False):
        """fma=True loads the FP32 build with FMA contraction (the rounding
        of the reference's own -O2 -mfma build) instead of the strict one."""
        self.dtype = np.dtype(dtype)
        name = "liboracle_f32.so" if self.dtype == np.float32 else "liboracle_f64.so"
        if fma:
            assert self.dtype == np.float32
            name = "liboracle_f32_fma.so"
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
        self.lib = ctypes.CDLL(path)
        self.real = ctypes.c_float if self.dtype == np.float32 else ctypes.c_double
        assert self.lib.oracle_real_size() == self.dtype.itemsize

    # -- helpers -----------------------------------------------------------
    def _p(self, a: np.ndarray):
        assert a.dtype == self.dtype and a.flags.c_contiguous, (a.dtype, a.flags)
        return a.ctypes.data_as(ctypes.c_void_p)

    def arr(self, a) -> np.ndarray:
        return np.ascontiguousarray(a, dtype=self.dtype)

    @staticmethod
    def _dims(q: np.ndarray):
        ny, nz, nx = q.shape
        return nx, ny, nz

    # -- restated reference routines -------------------------------------
    def compute_wind(self, dom, dt, advect_density):
        nx, ny, nz = dom.nx, dom.ny, dom.nz
        U = np.zeros(dom.shape, self.dtype)
        V = np.zeros(dom.shape, self.dtype)
        W = np.zeros(dom.shape, self.dtype)
        a = [self.arr(getattr(dom, n)) for n in
             ("u", "v", "w", "density", "jacobian_u", "jacobian_v", "jacobian_w")]
        self.lib.oracle_compute_wind(nx, ny, nz, self.real(dom.dx), self.real(dt),
                                     int(advect_density), *[self._p(x) for x in a],
                                     self._p(U), self._p(V), self._p(W))
        return U, V, W

    def upwind(self, qin, U, V, W, rho, dz, jaco):
        nx, ny, nz = self._dims(qin)
        q = np.empty_like(qin)
        self.lib.oracle_upwind(nx, ny, nz, self._p(qin), self._p(U), self._p(V), self._p(W),
                               self._p(rho), self._p(q), self._p(dz), self._p(jaco))
        return q

    def fluxes(self, q, U, V, Wd, G):
        nx, ny, nz = self._dims(q)
        u2, v2, w2 = (np.empty_like(q) for _ in range(3))
        self.lib.oracle_fluxes(nx, ny, nz, self._p(q), self._p(U), self._p(V), self._p(Wd),
                               self._p(u2), self._p(v2), self._p(w2), self._p(G))
        return u2, v2, w2

    def fct_core(self, q1, l, U2, flux_is_w):
        q1, l, U2 = self.arr(q1), self.arr(l), self.arr(U2).copy()
        n = q1.shape[0]
        assert l.shape == (n,) and U2.shape == (n - 1,) and n >= 3
        f = np.empty(n, self.dtype)
        self.lib.oracle_fct_core(n, self._p(q1), self._p(l), self._p(U2), int(flux_is_w),
                                 self._p(f))
        return U2

    def flux_limiter(self, q0, q1, u2, v2, w2):
        nx, ny, nz = self._dims(q0)
        u2, v2, w2 = u2.copy(), v2.copy(), w2.copy()
        self.lib.oracle_flux_limiter(nx, ny, nz, self._p(q0), self._p(q1), self._p(u2),
                                     self._p(v2), self._p(w2))
        return u2, v2, w2

    def advect3d(self, q, rho, jaco, dz, U, V, W, order, fct):
        nx, ny, nz = self._dims(q)
        q = q.copy()
        rc = self.lib.oracle_advect3d(nx, ny, nz, self._p(q), self._p(rho), self._p(jaco),
                                      self._p(dz), self._p(U), self._p(V), self._p(W),
                                      int(order), int(fct))
        assert rc == 0, rc
        return q

    def advect(self, dom, options, dt, species=None):
        """advection_driver.f90:59-126 on copies of the domain species (or
        the given dict); returns the dict of advanced species."""
        src = dom.species if species is None else species
        names = list(src.keys())
        qs = [self.arr(src[n]).copy() for n in names]
        ptrs = (ctypes.c_void_p * len(qs))(*[q.ctypes.data for q in qs])
        f = [self.arr(getattr(dom, n)) for n in
             ("u", "v", "w", "density", "jacobian", "jacobian_u", "jacobian_v", "jacobian_w",
              "advection_dz")]
        rc = self.lib.oracle_advect(
            dom.nx, dom.ny, dom.nz, self.real(dom.dx), self.real(dt),
            int(options.adv_options.mpdata_order), int(options.adv_options.MPDATA_FCT),
            int(options.parameters.advect_density), int(options.time_options.RK3),
            *[self._p(x) for x in f], len(qs), ptrs)
        assert rc == 0, rc
        return dict(zip(names, qs))

    def run(self, dom, options, dt, steps, bc_x=0, bc_y=0, species=None):
        """steps x (advect, then ring boundary conditions)."""
        cur = {n: self.arr(a).copy() for n, a in (species or dom.species).items()}
        for _ in range(steps):
            cur = self.advect(dom, options, dt, cur)
            for q in cur.values():
                self.apply_bc(q, bc_x, bc_y)
        return cur

    # -- standard scheme, adv=1 (std_oracle.c) ----------------------------
    @staticmethod
    def tile(dom):
        """The tile bounds its..ite, jts..jte as 0-based memory offsets."""
        return (dom.its - dom.ims, dom.ite - dom.ims, dom.jts - dom.jms, dom.jte - dom.jms)

    def std_compute_wind(self, dom, dt, advect_density):
        """adv_std_compute_wind (advect.f90:444-492): U_m on x faces
        (ny, nz, nx+1), V_m on y faces (ny+1, nz, nx), W_m and rho on cells."""
        Um = np.zeros(dom.field_shape("u"), self.dtype)
        Vm = np.zeros(dom.field_shape("v"), self.dtype)
        Wm = np.zeros(dom.shape, self.dtype)
        rho = np.zeros(dom.shape, self.dtype)
        a = [self.arr(getattr(dom, n)) for n in
             ("u", "v", "w", "density", "advection_dz", "jacobian_u", "jacobian_v",
              "jacobian_w")]
        self.lib.oracle_std_compute_wind(dom.nx, dom.ny, dom.nz, self.real(dom.dx),
                                         self.real(dt), int(advect_density),
                                         *[self._p(x) for x in a], self._p(Um), self._p(Vm),
                                         self._p(Wm), self._p(rho))
        return Um, Vm, Wm, rho

    def std_advect3d(self, dom, qf, qold, winds, horder, vorder, t_factor=1.0, flux_corr=0,
                     fc_bounds=None):
        """adv_std_advect3d (advect.f90:259-320): returns the new qfluxes.
        winds = std_compute_wind(...); fc_bounds defaults to flux_corr == 1."""
        Um, Vm, Wm, rho = winds
        qf = self.arr(qf).copy()
        qold = self.arr(qold)
        fcb = (flux_corr == 1) if fc_bounds is None else bool(fc_bounds)
        dz, jaco = self.arr(dom.advection_dz), self.arr(dom.jacobian)
        self.lib.oracle_std_advect3d(dom.nx, dom.ny, dom.nz, *self.tile(dom), int(horder),
                                     int(vorder), int(fcb), self._p(qf), self._p(qold),
                                     self._p(dz), self._p(jaco), self._p(Um), self._p(Vm),
                                     self._p(Wm), self._p(rho), self.real(t_factor),
                                     int(flux_corr))
        return qf

    def std_advect(self, dom, options, dt, species=None):
        """advection_driver.f90:59-126 with adv=1: the winds once, then each
        species by RK3 or one Euler step.  Returns the dict of new species."""
        ao = options.adv_options
        winds = self.std_compute_wind(dom, dt, options.parameters.advect_density)
        Um, Vm, Wm, rho = winds
        dz, jaco = self.arr(dom.advection_dz), self.arr(dom.jacobian)
        out = {}
        for n, a in (dom.species if species is None else species).items():
            q = self.arr(a).copy()
            self.lib.oracle_std_advect_species(
                dom.nx, dom.ny, dom.nz, *self.tile(dom), int(ao.h_order), int(ao.v_order),
                int(ao.flux_corr), int(options.time_options.RK3), self._p(q), self._p(dz),
                self._p(jaco), self._p(Um), self._p(Vm), self._p(Wm), self._p(rho))
            out[n] = q
        return out

    def std_flux3(self, q, u, v, w, horder, vorder, i_s, i_e, j_s, j_e):
        """flux3 (advect.f90:134-257) over cells i_s..i_e x j_s..j_e (0-based,
        inclusive); u on x faces, v on y faces, w on cells.  Returns fx
        (ny, nz, nx+1), fy (ny+1, nz, nx), fz (ny, nz+1, nx); faces outside
        the range are 0."""
        nx, ny, nz = self._dims(q)
        fx = np.zeros((ny, nz, nx + 1), self.dtype)
        fy = np.zeros((ny + 1, nz, nx), self.dtype)
        fz = np.zeros((ny, nz + 1, nx), self.dtype)
        q, u, v, w = (self.arr(a) for a in (q, u, v, w))
        self.lib.oracle_std_flux3(nx, ny, nz, i_s, i_e, j_s, j_e, int(horder), int(vorder),
                                  self._p(q), self._p(u), self._p(v), self._p(w), self._p(fx),
                                  self._p(fy), self._p(fz))
        return fx, fy, fz

    def std_flux_corr(self, q, u, v, w, fx, fy, fz, jaco, dz, rho, its, ite, jts, jte):
        """WRF_flux_corr (adv_fluxcorr.f90:37-140) of the given fluxes; returns
        the corrected copies."""
        nx, ny, nz = self._dims(q)
        fx, fy, fz = (self.arr(a).copy() for a in (fx, fy, fz))
        a = [self.arr(x) for x in (q, u, v, w)]
        b = [self.arr(x) for x in (jaco, dz, rho)]
        self.lib.oracle_std_flux_corr(nx, ny, nz, its, ite, jts, jte, *[self._p(x) for x in a],
                                      self._p(fx), self._p(fy), self._p(fz),
                                      *[self._p(x) for x in b])
        return fx, fy, fz

    # -- wind balance and time step (dyn_oracle.c) ------------------------
    def _winds(self, dom):
        return [self.arr(getattr(dom, n)) for n in
                ("u", "v", "density", "jacobian_u", "jacobian_v", "jacobian_w", "advection_dz")]

    def calc_divergence(self, dom, advect_density, w=None, horz_only=True):
        """calc_divergence (wind.f90:201-246) of the domain's winds."""
        u, v, rho, ju, jv, jw, dz = self._winds(dom)
        w = self.arr(dom.w if w is None else w)
        div = np.zeros(dom.shape, self.dtype)
        self.lib.oracle_calc_divergence(dom.nx, dom.ny, dom.nz, self.real(dom.dx),
                                        int(advect_density), int(horz_only), self._p(u),
                                        self._p(v), self._p(w), self._p(rho), self._p(ju),
                                        self._p(jv), self._p(jw), self._p(dz), self._p(div))
        return div

    def balance_w(self, dom, advect_density, i_s=None, i_e=None, j_s=None, j_e=None):
        """balance_uvw (wind.f90:88-164, update=.false.): w from the
        horizontal divergence over columns i_s..i_e x j_s..j_e (0-based,
        inclusive; default: every column), zero elsewhere."""
        i_s = 0 if i_s is None else i_s
        j_s = 0 if j_s is None else j_s
        i_e = dom.nx - 1 if i_e is None else i_e
        j_e = dom.ny - 1 if j_e is None else j_e
        u, v, rho, ju, jv, jw, dz = self._winds(dom)
        div = self.calc_divergence(dom, advect_density)
        w = np.zeros(dom.shape, self.dtype)
        self.lib.oracle_calc_w(dom.nx, dom.ny, dom.nz, i_s, i_e, j_s, j_e, int(advect_density),
                               self._p(div), self._p(dz), self._p(jw), self._p(rho), self._p(w))
        return w

    def compute_dt(self, dom, dz_levels, cfl, strictness, its=None, ite=None, jts=None,
                   jte=None, w=None):
        """compute_dt (time_step.f90:51-168, use_density=.false.) over the
        tile its..ite x jts..jte (0-based, inclusive; default: all cells)."""
        its = 0 if its is None else its
        jts = 0 if jts is None else jts
        ite = dom.nx - 1 if ite is None else ite
        jte = dom.ny - 1 if jte is None else jte
        u, v = self.arr(dom.u), self.arr(dom.v)
        w = self.arr(dom.w if w is None else w)
        dzl = self.arr(dz_levels)
        assert dzl.shape == (dom.nz,)
        self.lib.oracle_compute_dt.restype = self.real
        return self.lib.oracle_compute_dt(dom.nx, dom.ny, dom.nz, its, ite, jts, jte,
                                          self.real(dom.dx), self._p(u), self._p(v), self._p(w),
                                          self._p(dzl), self.real(cfl), int(strictness))

    # -- lateral forcing (dyn_oracle.c) ------------------------------------
    def relax_filter(self, nx, ny, hs, west, east, north, south):
        """init_relax_filters (domain_obj.f90:3285-3390): the (ny, nx) filter."""
        f = np.zeros((ny, nx), self.dtype)
        self.lib.oracle_relax_filter(nx, ny, hs, int(west), int(east), int(north), int(south),
                                     self._p(f))
        return f

    def apply_forcing(self, q, fhi, dqdt, filt, dt):
        """apply_forcing (domain_obj.f90:3493-3503) for one 3-D boundary-forced
        variable: returns the new (q, forcing_hi)."""
        nx, ny, nz = self._dims(q)
        q, fhi = self.arr(q).copy(), self.arr(fhi).copy()
        dqdt, filt = self.arr(dqdt), self.arr(filt)
        assert filt.shape == (ny, nx)
        self.lib.oracle_apply_forcing(nx, ny, nz, ctypes.c_double(dt), self._p(filt), self._p(q),
                                      self._p(fhi), self._p(dqdt))
        return q, fhi

    def divergence(self, U, V, W, dz):
        nx, ny, nz = self._dims(U)
        mx = ctypes.c_double()
        self.lib.oracle_divergence.restype = ctypes.c_long
        n = self.lib.oracle_divergence(nx, ny, nz, self._p(U), self._p(V), self._p(W),
                                       self._p(dz), ctypes.byref(mx))
        return int(n), mx.value

    def mass(self, q, jaco, rho, dz) -> float:
        nx, ny, nz = self._dims(q)
        self.lib.oracle_mass.restype = ctypes.c_double
        return self.lib.oracle_mass(nx, ny, nz, self._p(q), self._p(jaco), self._p(rho),
                                    self._p(dz))

    def apply_bc(self, q, bc_x, bc_y):
        nx, ny, nz = self._dims(q)
        self.lib.oracle_apply_bc(nx, ny, nz, self._p(q), int(bc_x), int(bc_y))
        return q
