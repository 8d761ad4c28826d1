"""How far the FP32 arithmetic of the reference is from its FP64 build, and
where a different rounding order moves it, on a 130 x 70 x 30 domain of the
10-species workload (fractal terrain, 15 m/s westerly, w from continuity).

Every field stays within 1e-6 of the FP64 oracle except theta with FCT on:
there the difference sits in the top levels and grows about 8-fold per step
at mpdata_order 3 (1.8e-5 after 2 steps, 3.4e-4 after 4, 8.6e-4 after 10,
relative to max |theta|).  The model-top cell exchanges theta * W through
the lid without upwinding (adv_mpdata.f90:96-97), a flux the limiter's
bounds for that cell do not include (the w-column end of
adv_mpdata_FCT_core.f90:92-99, w2(kme) = 0 at adv_mpdata.f90:318), so the top
level moves by kelvins per step while its limiter factors are ratios of
differences of ~370 K values (FCT_core:103-115, the third pass taking its
bounds from q0 = q1, adv_mpdata.f90:388-397).  Last-bit differences there
feed back step after step.  It is a property of the reference's FP32
arithmetic, not of the GPU kernels: the GPU path equals the FP32 oracle bit for
bit.  An FP32 build with FMA contraction moves the same cells the same way
and nothing else, which is why the kernels keep the reference's operation
order instead of tolerating reordering.
"""
import numpy as np
import pytest

from hicar_amd import synthetic

TOP = 5  # the top levels where theta's difference lives


def case(order, fct):
    c = synthetic.config2(nx=130, ny=70, nz=30, steps=10)
    c.options.adv_options.mpdata_order = order
    c.options.adv_options.MPDATA_FCT = fct
    return c


def run(orc, c, steps):
    d = c.domain
    return orc.run(d, c.options, c.dt, steps, c.bc_x, c.bc_y,
                   {n: a.copy() for n, a in d.species.items()})


def rel(a, b, levels=slice(None)):
    """max |a - b| over the levels, relative to max |b| over the field"""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a[:, levels] - b[:, levels]).max() / np.abs(b).max())


def test_fp32_against_fp64_theta_diverges_at_the_top_with_fct(oracle32, oracle64):
    c = case(3, True)
    a, b = run(oracle32, c, 4), run(oracle64, c, 4)
    below = slice(0, c.domain.nz - TOP)
    assert rel(a["theta"], b["theta"]) > 1e-4  # beyond the 1e-4 bar after 4 steps
    assert rel(a["theta"], b["theta"], below) < 1e-5  # the rest of the column is not
    for n in a:
        if n != "theta":
            assert rel(a[n], b[n]) < 1e-6, n
    # the difference grows step after step: 2 steps are still below the bar
    a2, b2 = run(oracle32, c, 2), run(oracle64, c, 2)
    assert 1e-6 < rel(a2["theta"], b2["theta"]) < 1e-4


def test_fp32_against_fp64_without_fct_stays_at_rounding(oracle32, oracle64):
    c = case(3, False)
    a, b = run(oracle32, c, 4), run(oracle64, c, 4)
    for n in a:
        assert rel(a[n], b[n]) < 2e-6, n


def test_fma_contraction_moves_theta_at_the_top_only(oracle32, oracle32_fma):
    """The FP32 oracle built with FMA contraction (oracle/Makefile's third
    build) against the plain FP32 one, order 3 + FCT, 5 steps: theta's top
    levels move by about 6e-4, the hydrometeors by less than 1e-7, and below
    the top levels nothing moves."""
    c = case(3, True)
    a, f = run(oracle32, c, 5), run(oracle32_fma, c, 5)
    below = slice(0, c.domain.nz - TOP)
    assert rel(f["theta"], a["theta"]) > 1e-4
    for n in a:
        assert rel(f[n], a[n], below) < 1e-9, n
        if n != "theta":
            assert rel(f[n], a[n]) < 1e-7, n
