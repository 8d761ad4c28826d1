"""Grid shapes at the edges of the kernels' tiling, against the FP32 oracle
bit for bit, for the exact and the fused kernels: the smallest grid the
library accepts (3x3x3, one interior cell), and widths and heights just
below, at and above the corrective kernel's 60x16 output tile and 64x20
staged tile, with orders 1-3, FCT on and off."""
import numpy as np
import pytest

from test_gpu_parity import fresh_species, gpu_run, random_domain

pytestmark = pytest.mark.gpu

SHAPES = [(3, 3, 3), (4, 5, 3), (59, 15, 4), (60, 16, 4), (61, 17, 5), (64, 20, 3),
          (65, 21, 6), (120, 32, 4), (124, 33, 7)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("order,fct", [(1, False), (2, True), (3, True), (3, False)])
@pytest.mark.parametrize("variant", [0, 1])
def test_tile_edge_shapes_match_oracle(oracle32, shape, order, fct, variant):
    nx, ny, nz = shape
    case = random_domain(nx, ny, nz, 7000 + nx * 31 + ny * 7 + nz, order, fct, True, nspec=2)
    ref = oracle32.run(case.domain, case.options, case.dt, case.steps, 0, 0, fresh_species(case))
    got = gpu_run(case, variant=variant)
    for n in ref:
        np.testing.assert_array_equal(got[n].view(np.uint32), ref[n].view(np.uint32), err_msg=n)


# tile (interior) sizes: one cell, and widths and heights around the 63x7 tile
STD_TILES = [(1, 1, 3), (4, 1, 4), (62, 6, 4), (63, 7, 5), (64, 8, 3), (69, 13, 4),
             (126, 14, 6), (127, 15, 3)]


@pytest.mark.parametrize("tile", STD_TILES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("h,v,fc,rk3", [(5, 3, 1, True), (3, 3, 0, False), (1, 1, 0, False)])
def test_std_tile_edge_shapes_match_oracle(oracle32, tile, h, v, fc, rk3):
    """The standard scheme (adv=1) with tiles of one cell and of widths and
    heights around its 63x7 launch tiles (the grid adds the halo the order
    and limiter need: 4 cells for the fifth-order flux with the limiter),
    two steps against the FP32 oracle bit for bit."""
    from hicar_amd import synthetic
    from hicar_amd.advection import Advection
    wi, hj, nz = tile

    def make(nx, ny):
        return synthetic.config_std(nx=nx, ny=ny, nz=nz, steps=2, h_order=h, v_order=v,
                                    flux_corr=fc, rk3=rk3, cfl=1.0 if rk3 else 0.4)
    halo = make(16, 16).domain.halo
    case = make(wi + 2 * halo, hj + 2 * halo)
    d = case.domain
    assert (d.ite - d.its + 1, d.jte - d.jts + 1) == (wi, hj)
    cur = {n: a.copy() for n, a in d.species.items()}
    for _ in range(2):
        cur = oracle32.std_advect(d, case.options, case.dt, cur)
    st = Advection(d, case.options)
    try:
        st.advect(case.dt)
        st.advect(case.dt)
        st.download()
    finally:
        st.close()
    for n in cur:
        np.testing.assert_array_equal(d.species[n].view(np.uint32), cur[n].view(np.uint32),
                                      err_msg=n)
