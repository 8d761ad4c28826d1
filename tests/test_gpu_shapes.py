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
