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


@pytest.mark.parametrize("variant", [0, 1])
def test_fortran_driver_matches_oracle(tmp_path, oracle32, variant):
    assert os.path.exists(EXE), "build the Fortran driver first (make -C fortran)"
    case = synthetic.config2(nx=40, ny=24, nz=10, steps=4, nspecies=3)
    d, o = case.domain, case.options
    ref = oracle32.run(d, o, case.dt, case.steps, case.bc_x, case.bc_y, fresh_species(case))
    hdr = np.array([d.nx, d.ny, d.nz, 3, 1, 7, len(d.species), o.adv_options.mpdata_order,
                    int(o.adv_options.MPDATA_FCT), int(o.parameters.advect_density), case.steps,
                    case.bc_x, case.bc_y, variant], np.int32)
    fin = tmp_path / "in.bin"
    fout = tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([d.dx, case.dt], np.float32).tobytes())
        for n in ("u", "v", "w", "density", "jacobian", "jacobian_u", "jacobian_v", "jacobian_w",
                  "advection_dz"):
            f.write(fortran_order(getattr(d, n)))
        for a in d.species.values():
            f.write(fortran_order(a))
    subprocess.run([EXE, str(fin), str(fout)], check=True, timeout=120)
    out = np.fromfile(fout, np.float32).reshape(len(d.species), d.ny, d.nz, d.nx)
    for s, n in enumerate(d.species):
        np.testing.assert_array_equal(out[s], ref[n], err_msg=n)
