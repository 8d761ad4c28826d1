"""The drop-in boundary from the Fortran side: fortran/test_icar_adv (built
with flang against fortran/icar_adv_c.f90) allocates ICAR-shaped arrays with
non-trivial lower bounds, binds them with c_loc, and drives the library
through the iso_c_binding interface; its output equals the oracle bit for
bit."""
import os
import subprocess

import numpy as np
import pytest

from hicar_amd import synthetic

from test_gpu_parity import fresh_species

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fortran", "test_icar_adv")


def fortran_order(a):
    # numpy (ny, nz, nx) C order is Fortran (nx, nz, ny) column-major
    return np.ascontiguousarray(a, dtype=np.float32).tobytes()


def run_driver(tmp_path, case, hdr):
    d = case.domain
    fin = tmp_path / "in.bin"
    fout = tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.array(hdr, np.int32).tobytes())
        f.write(np.array([d.dx, case.dt], np.float32).tobytes())
        for n in ("u", "v", "w", "density", "jacobian", "jacobian_u", "jacobian_v", "jacobian_w",
                  "advection_dz"):
            f.write(fortran_order(getattr(d, n)))
        for a in d.species.values():
            f.write(fortran_order(a))
    subprocess.run([EXE, str(fin), str(fout)], check=True, timeout=120)
    return np.fromfile(fout, np.float32).reshape(len(d.species), d.ny, d.nz, d.nx)


@pytest.mark.parametrize("variant", [0, 1])
def test_fortran_driver_matches_oracle(tmp_path, oracle32, variant):
    assert os.path.exists(EXE), "build the Fortran driver first (make -C fortran)"
    case = synthetic.config2(nx=40, ny=24, nz=10, steps=4, nspecies=3)
    d, o = case.domain, case.options
    ref = oracle32.run(d, o, case.dt, case.steps, case.bc_x, case.bc_y, fresh_species(case))
    hdr = [d.nx, d.ny, d.nz, 3, 1, 7, len(d.species), o.adv_options.mpdata_order,
           int(o.adv_options.MPDATA_FCT), int(o.parameters.advect_density), case.steps,
           case.bc_x, case.bc_y, variant, 2, 2, 1, 1, 0, 0]
    out = run_driver(tmp_path, case, hdr)
    for s, n in enumerate(d.species):
        np.testing.assert_array_equal(out[s], ref[n], err_msg=n)


def test_fortran_driver_standard_scheme(tmp_path, oracle32):
    """adv=1 with the run/complete_icar_options.nml settings (h_order 5,
    v_order 3, WRF limiter, RK3) and ICAR's halo for them, from Fortran."""
    assert os.path.exists(EXE), "build the Fortran driver first (make -C fortran)"
    case = synthetic.config_std(nx=40, ny=30, nz=9, steps=2, nspecies=3, cfl=1.0)
    d, o = case.domain, case.options
    ref = {n: a.copy() for n, a in d.species.items()}
    for _ in range(case.steps):
        ref = oracle32.std_advect(d, o, case.dt, ref)
    ao = o.adv_options
    hdr = [d.nx, d.ny, d.nz, 1, 1, 1, len(d.species), 2, 1, int(o.parameters.advect_density),
           case.steps, 0, 0, 0, d.halo, 1, ao.h_order, ao.v_order, ao.flux_corr,
           int(o.time_options.RK3)]
    out = run_driver(tmp_path, case, hdr)
    for s, n in enumerate(d.species):
        np.testing.assert_array_equal(out[s], ref[n], err_msg=n)
