"""The 48-domain ensemble at its stated size: nx, ny drawn from 64..512,
nz from {20, 30, 40} (97 M cells in all), 6 FP32 species plus a constant
tracer, mpdata_order 2 with FCT, each domain with its own fractal terrain and
winds.  ICAR holds one domain per process (adv_mpdata.f90:434-440), so every
member of the ensemble must step exactly as it does on its own:
  - one launch per pass over all 48 domains, eager and replayed from a HIP
    graph, equals each member's own single-handle run bit for bit;
  - the three smallest members equal the FP32 oracle at full size;
  - the constant tracer stays constant to 1e-5 (the synthetic w balances the
    scheme's Courant numbers) and, with FCT, no species goes negative.
"""
import time

import numpy as np
import pytest

from hicar_amd import synthetic
from hicar_amd.advection import Ensemble, MpdataAdvection

from test_gpu_parity import fresh_species

pytestmark = pytest.mark.gpu

STEPS = 4


def restore(case, values):
    for n, a in values.items():
        case.domain.species[n][...] = a


def test_c3_full_size_ensemble(oracle32):
    t0 = time.perf_counter()
    cases = synthetic.config3(48, steps=STEPS)
    for c in cases:
        d, o = c.domain, c.options
        assert 64 <= d.nx <= 512 and 64 <= d.ny <= 512 and d.nz in (20, 30, 40)
        assert o.adv_options.mpdata_order == 2 and o.adv_options.MPDATA_FCT
        o.vars_to_advect = list(o.vars_to_advect) + ["qtr"]
        d.set_var_lists(o, {**d.species, "qtr": np.full(d.shape, 2.5, np.float32)})
    init = [fresh_species(c) for c in cases]
    cells = sum(c.domain.nx * c.domain.ny * c.domain.nz for c in cases)
    print(f"\nC3: 48 domains, {cells / 1e6:.1f} M cells, generated in "
          f"{time.perf_counter() - t0:.0f} s", flush=True)

    singles = []
    for c, ini in zip(cases, init):
        restore(c, ini)
        st = MpdataAdvection(c.domain, c.options)
        try:
            st.run(c.dt, STEPS, c.bc_x, c.bc_y)
            st.download()
        finally:
            st.close()
        singles.append(fresh_species(c))
    print(f"C3: 48 single-domain runs done at {time.perf_counter() - t0:.0f} s", flush=True)

    bcs = {(c.bc_x, c.bc_y) for c in cases}
    assert len(bcs) == 1
    bc_x, bc_y = bcs.pop()
    dts = [c.dt for c in cases]
    for graph in (False, True):
        for c, ini in zip(cases, init):
            restore(c, ini)
        members = [MpdataAdvection(c.domain, c.options) for c in cases]
        try:
            ens = Ensemble(members)
            try:
                if graph:  # recorded on the first call, replayed on the second
                    ens.run(dts, STEPS // 2, bc_x, bc_y, graph=True)
                    ens.run(dts, STEPS - STEPS // 2, bc_x, bc_y, graph=True)
                else:
                    ens.run(dts, STEPS, bc_x, bc_y)
                ens.sync()
            finally:
                ens.close()
            for m in members:
                m.download()
        finally:
            for m in members:
                m.close()
        for c, ref in zip(cases, singles):
            for n, a in ref.items():
                np.testing.assert_array_equal(c.domain.species[n].view(np.uint32),
                                              a.view(np.uint32),
                                              err_msg=f"{c.name} {n} graph={graph}")
        print(f"C3: ensemble (graph={graph}) equal to the single runs at "
              f"{time.perf_counter() - t0:.0f} s", flush=True)

    for c, got in zip(cases, singles):
        for n, a in got.items():
            assert np.isfinite(a).all() and a.min() >= 0.0, (c.name, n)
        assert np.abs(got["qtr"].astype(np.float64) / 2.5 - 1).max() < 1e-5, c.name

    order = sorted(range(len(cases)), key=lambda i: cases[i].domain.nx * cases[i].domain.ny *
                   cases[i].domain.nz)
    for i in order[:3]:
        c = cases[i]
        ref = oracle32.run(c.domain, c.options, c.dt, STEPS, c.bc_x, c.bc_y, init[i])
        for n, a in ref.items():
            np.testing.assert_array_equal(singles[i][n].view(np.uint32), a.view(np.uint32),
                                          err_msg=f"{c.name} {n}")
    print(f"C3: three smallest members equal to the oracle at {time.perf_counter() - t0:.0f} s",
          flush=True)
