"""The drop-in boundary from the Fortran side: fortran/test_icar_adv (built
with flang against fortran/icar_adv_c.f90) allocates ICAR-shaped arrays with
non-trivial lower bounds, binds them with c_loc, and drives the library
through the iso_c_binding interface; its output equals the oracle bit for
bit."""
import os
import subprocess

import numpy as np
import pytest

from hicar_amd import _lib, synthetic
from hicar_amd.advection import Advection

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


EXT = os.path.join(ROOT, "fortran", "test_icar_adv_ext")


def windy(case, seed):
    """Winds that balance_uvw has work to do on, and species with negative
    cells for enforce_limits."""
    d = case.domain
    r = np.random.default_rng(seed)
    d.v[...] = r.uniform(-4, 4, d.v.shape)
    d.u[...] += r.uniform(-1, 1, d.u.shape).astype(np.float32)
    for a in d.species.values():
        a[...] -= np.float32(0.05) * a.max()
    return r


@pytest.mark.parametrize("scheme", ["mpdata", "std"])
def test_fortran_entry_points_around_the_step(tmp_path, monkeypatch, scheme):
    """fortran/test_icar_adv_ext calls balance_w, compute_dt, run,
    max_courant, set_relax_filter / set_forcing / apply_forcing,
    enforce_limits, mass_budget, halo_bytes, launch_config and snapshot /
    snapshot_wait through icar_adv_c.f90; the same calls through the Python
    binding (whose results the oracle tests pin) give the same bits."""
    assert os.path.exists(EXT), "build the Fortran drivers first (make -C fortran)"
    # the autotuner's choice follows timings that differ between processes:
    # both sides take the same fixed configuration
    monkeypatch.setenv("ICAR_ADV_CORR_CFG", "3")
    monkeypatch.setenv("ICAR_ADV_KSPLIT", "2")
    if scheme == "std":
        case = synthetic.config_std(nx=40, ny=30, nz=9, steps=2, nspecies=3, cfl=1.0)
        variant = 0
    else:
        case = synthetic.config2(nx=40, ny=24, nz=10, steps=2, nspecies=3)
        variant = 1
    d, o = case.domain, case.options
    o.kernel_variant = variant
    r = windy(case, 31)
    ao = o.adv_options
    cfl, strictness, flags, fdt = 0.9, 3, (1, 0, 1, 1), 37.5
    dzl = d.advection_dz[0, :, 0].astype(np.float32)
    fhi = r.uniform(0.5, 2.0, d.shape).astype(np.float32)
    dqdt = r.uniform(-1e-4, 1e-4, d.shape).astype(np.float32)
    w0 = d.w.copy()
    species0 = {n: a.copy() for n, a in d.species.items()}

    hdr = [d.nx, d.ny, d.nz, 1 - d.halo, 1, 1 - d.halo, len(d.species), ao.mpdata_order,
           int(ao.MPDATA_FCT), int(o.parameters.advect_density), case.steps, case.bc_x,
           case.bc_y, variant, d.halo, int(o.physics.advection), ao.h_order, ao.v_order,
           ao.flux_corr, int(o.time_options.RK3)]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.array(hdr, np.int32).tobytes())
        f.write(np.array([d.dx, case.dt], np.float32).tobytes())
        for n in ("u", "v", "w", "density", "jacobian", "jacobian_u", "jacobian_v", "jacobian_w",
                  "advection_dz"):
            f.write(fortran_order(getattr(d, n)))
        for a in d.species.values():
            f.write(fortran_order(a))
        f.write(np.float32(cfl).tobytes() + np.array([strictness, *flags], np.int32).tobytes())
        f.write(np.float64(fdt).tobytes() + dzl.tobytes())
        f.write(fortran_order(fhi) + fortran_order(dqdt))
    subprocess.run([EXT, str(fin), str(fout)], check=True, timeout=120)
    raw = fout.read_bytes()
    ns, cells = len(d.species), d.nx * d.ny * d.nz
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype, count, off)
        off += a.nbytes
        return a

    f_w = take(np.float32, cells).reshape(d.shape)
    f_q = take(np.float32, ns * cells).reshape(ns, *d.shape)
    f_courant = take(np.float64, 3)
    f_dt = take(np.float32, 1)[0]
    f_budget = take(np.float64, ns)
    f_bytes = take(np.int64, 4)
    f_cfg = take(np.int32, 2)
    assert off == len(raw)

    st = Advection(d, o)  # the same inputs, through the Python binding
    try:
        st.balance_w()
        dt = st.compute_dt(cfl, strictness, dzl)
        st.run(case.dt, case.steps, case.bc_x, case.bc_y)
        courant = st.max_courant()
        st.set_relax_filter(*flags, halo=d.halo)
        st.set_forcing(st.names[0], fhi, dqdt)
        st.apply_forcing(fdt)
        st.enforce_limits()
        budget = st.mass_budget()
        nbytes = [st.halo_bytes(s) for s in (_lib.NORTH, _lib.SOUTH, _lib.EAST, _lib.WEST)]
        cfg = st.launch_config()
        st.snapshot(species=True, fields=True)
        st.snapshot_wait()
    finally:
        st.close()
    assert not np.array_equal(d.w, w0)  # balance_w changed w
    np.testing.assert_array_equal(f_w.view(np.uint32), d.w.view(np.uint32))
    for s, n in enumerate(d.species):
        np.testing.assert_array_equal(f_q[s].view(np.uint32), d.species[n].view(np.uint32),
                                      err_msg=n)
        assert d.species[n].min() >= 0, n  # enforce_limits ran last
    assert not np.array_equal(d.species[st.names[0]], species0[st.names[0]])
    assert np.float32(f_dt) == np.float32(dt) and dt > 0.1
    np.testing.assert_array_equal(f_courant, courant)
    np.testing.assert_array_equal(f_budget, budget)
    assert list(f_bytes) == nbytes and min(nbytes) > 0
    assert tuple(int(c) for c in f_cfg) == cfg
    if scheme == "std":
        assert cfg == (-1, -1)
    else:
        assert cfg == (3, 2)
