"""The large single domain, batched and mixed precision: 4096 x 4096 x 50
(839 M cells), the 12 species of the precision case with the hydrometeor
masses and numbers stored in BF16 with stochastic rounding and the tracer in
BF16 rounded to nearest (per-species power-of-two scaling), mpdata_order 2 with FCT, 20 steps at Courant 0.5.  About 180 GB
of device memory.  Too large for the oracle; checked through properties that
do not depend on its size:
  - every value stays finite;
  - with FCT no species goes negative (all start non-negative);
  - a constant tracer stays constant: the synthetic w balances the scheme's
    Courant numbers, so MPDATA moves nothing but rounding, and the BF16
    value 2.5 is exact; its FP64 mass budget does not move either.
"""
import time

import numpy as np
import pytest

from hicar_amd import _lib, synthetic
from hicar_amd.advection import Advection, MpdataAdvection

pytestmark = pytest.mark.gpu

STEPS = 20


def test_c5_full_size_mixed_precision_properties():
    t0 = time.perf_counter()
    case = synthetic.config5(steps=STEPS)
    d = case.domain
    assert d.nx * d.ny * d.nz == 838_860_800 and len(d.species) == 12
    d.species["qtr"][...] = 2.5
    # round to nearest: a value within half an ulp of 2.5 is stored as 2.5
    # (stochastic rounding would move a few cells by an ulp)
    kind = _lib.STORE_BF16
    d.storage["qtr"] = (kind, synthetic.pow2_scale(d.species["qtr"], kind))
    assert d.storage["qtr"][1] == 2  # 2.5 = 0.625 * 2^2, exact in BF16
    t1 = time.perf_counter()
    st = MpdataAdvection(d, case.options)
    try:
        m0 = st.mass_budget()
        st.run(case.dt, STEPS, case.bc_x, case.bc_y)
        m1 = st.mass_budget()
        st.download()
    finally:
        st.close()
    t2 = time.perf_counter()
    print(f"C5: generated in {t1 - t0:.0f} s, uploaded and {STEPS} steps in {t2 - t1:.0f} s")
    names = list(d.species)
    for n, a in d.species.items():
        assert np.isfinite(a).all(), n
        assert a.min() >= 0.0, (n, float(a.min()))
    c = d.species["qtr"].astype(np.float64)
    assert np.abs(c / 2.5 - 1).max() < 1e-5
    k = names.index("qtr")
    assert abs(m1[k] / m0[k] - 1) < 1e-6, (m0[k], m1[k])


def test_standard_scheme_past_2_29_cells():
    """h_order 5, v_order 3, the WRF limiter and RK3, two steps: finite; a
    species that starts non-negative stays so below the 2s-1 = 3 levels the
    unlimited model-top flux reaches (up to FP32 rounding of nearly equal
    fluxes); a constant tracer stays constant (w balances this scheme's
    mass fluxes)."""
    case = synthetic.config5(nx=4096, ny=4096, nz=33, steps=2, scheme="std", names=("qc", "qtr"))
    d = case.domain
    assert (d.nx + 2) * (d.ny + 2) * d.nz > 2 ** 29
    d.species["qtr"][...] = 2.5
    st = Advection(d, case.options)
    try:
        st.advect(case.dt)
        st.advect(case.dt)
        st.download()
    finally:
        st.close()
    q = d.species["qc"]
    assert np.isfinite(q).all()
    assert q[:, : d.nz - 3].min() >= -1e-7 * q.max()
    c = d.species["qtr"].astype(np.float64)
    assert np.abs(c / 2.5 - 1).max() < 1e-5
