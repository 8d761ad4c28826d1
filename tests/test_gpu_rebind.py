"""Binding species after a run was recorded in a HIP graph, or after an
ensemble was built: a binding may move the species' device buffer (a new
storage format) and the fused path rebuilds its tables, so a recorded run is
dropped and recorded again, and an ensemble refuses to run until it is built
again."""
import copy
import ctypes

import numpy as np
import pytest

from hicar_amd import _lib, synthetic
from hicar_amd.advection import Advection, Ensemble
from hicar_amd.domain import Domain

pytestmark = pytest.mark.gpu

_fp = ctypes.POINTER(ctypes.c_float)


def bind(st, idx, arr, storage=_lib.STORE_F32, e=0):
    """icar_adv_bind_species on an existing handle (the Python class binds
    only at construction)."""
    st._check(st.lib.icar_adv_bind_species(st.h, idx, arr.ctypes.data_as(_fp), storage, e))
    st._keep.append(arr)


def small_case():
    case = synthetic.config2(nx=40, ny=30, nz=8, steps=2, nspecies=3)
    case.options.kernel_variant = 1
    return case


def test_binding_after_a_recorded_run_records_it_again():
    case = small_case()
    d, o = case.domain, case.options
    names = list(d.species)
    r = np.random.default_rng(23)
    tracer = r.uniform(0.5, 1.5, d.shape).astype(np.float32)
    tracer0 = tracer.copy()
    st = Advection(d, o)
    try:
        st.set_graph(True)
        st.run(case.dt, 2)  # recorded with three FP32 species
        st.download()
        # species 1 moves to BF16 storage (a new device buffer), a fourth is added
        e1 = synthetic.pow2_scale(d.species[names[1]], _lib.STORE_BF16)
        bind(st, 1, d.species[names[1]], _lib.STORE_BF16, e1)
        bind(st, 3, tracer)
        st.upload(fields=False)
        # the same host state, bound that way from the start, run eagerly
        d2 = Domain(d.nx, d.ny, d.nz, d.dx, halo=d.halo)
        for n in d.FIELDS:
            getattr(d2, n)[...] = getattr(d, n)
        o2 = copy.deepcopy(o)
        o2.vars_to_advect = names + ["tracer"]
        d2.set_var_lists(o2, {**{n: a.copy() for n, a in d.species.items()},
                              "tracer": tracer.copy()})
        d2.storage[names[1]] = (_lib.STORE_BF16, e1)
        assert list(d2.species) == names + ["tracer"]
        st.run(case.dt, 2)  # same dt and steps: a stale recording would be replayed here
        st.download()
        ref = Advection(d2, o2)
        try:
            ref.run(case.dt, 2)
            ref.download()
        finally:
            ref.close()
    finally:
        st.close()
    for n in names:
        np.testing.assert_array_equal(d.species[n].view(np.uint32),
                                      d2.species[n].view(np.uint32), err_msg=n)
    np.testing.assert_array_equal(tracer.view(np.uint32), d2.species["tracer"].view(np.uint32))
    assert not np.array_equal(tracer, tracer0)  # the added species was advanced


def test_ensemble_refuses_to_run_after_a_member_binds_again():
    cases = [small_case(), synthetic.config2(nx=30, ny=26, nz=6, steps=2, nspecies=3)]
    cases[1].options.kernel_variant = 1
    members = [Advection(c.domain, c.options) for c in cases]
    try:
        ens = Ensemble(members)
        try:
            dts = [c.dt for c in cases]
            ens.run(dts, 1)
            ens.sync()
            d = cases[0].domain
            name = list(d.species)[1]
            members[0].download()
            bind(members[0], 1, d.species[name], _lib.STORE_BF16,
                 synthetic.pow2_scale(d.species[name], _lib.STORE_BF16))
            with pytest.raises(_lib.IcarAdvError) as err:
                ens.run(dts, 1)
            assert err.value.code == _lib.EINVAL and "bound a species" in str(err.value)
        finally:
            ens.close()
        members[0].upload(fields=False)
        ens = Ensemble(members)  # built again over the new buffers
        try:
            ens.run(dts, 1, graph=True)
            ens.sync()
        finally:
            ens.close()
        for m in members:
            m.download()
        for c in cases:
            for n, a in c.domain.species.items():
                assert np.isfinite(a).all(), n
    finally:
        for m in members:
            m.close()


def test_forcing_refused_after_a_forced_species_moves_to_16_bit():
    """Lateral forcing updates FP32 species; a forced species bound again in
    BF16 makes icar_adv_apply_forcing return EINVAL, before any launch (the
    other species' forcing is not applied either), and binding it back to
    FP32 lets the forcing run again."""
    case = small_case()
    d, o = case.domain, case.options
    names = list(d.species)
    r = np.random.default_rng(29)
    f = r.uniform(0.5, 2.0, d.shape).astype(np.float32)
    t = r.uniform(-1e-4, 1e-4, d.shape).astype(np.float32)
    st = Advection(d, o)
    try:
        st.set_relax_filter(1, 1, 1, 1, halo=2)
        st.set_forcing(names[0], f, t)
        st.set_forcing(names[2], f, t)
        st.apply_forcing(10.0)
        st.download()
        before = {n: a.copy() for n, a in d.species.items()}
        bind(st, 0, d.species[names[0]], _lib.STORE_BF16,
             synthetic.pow2_scale(d.species[names[0]], _lib.STORE_BF16))
        st.upload(fields=False)
        st.download()
        held = {n: a.copy() for n, a in d.species.items()}
        with pytest.raises(_lib.IcarAdvError) as err:
            st.apply_forcing(10.0)
        assert err.value.code == _lib.EINVAL and "16-bit" in str(err.value)
        st.download()
        for n in names:  # nothing ran
            np.testing.assert_array_equal(d.species[n], held[n], err_msg=n)
        bind(st, 0, d.species[names[0]])  # FP32 again
        st.upload(fields=False)
        st.apply_forcing(10.0)
        st.download()
        assert not np.array_equal(d.species[names[2]], before[names[2]])
    finally:
        st.close()


@pytest.mark.parametrize("storage", [_lib.STORE_F32, _lib.STORE_BF16])
def test_binding_at_order_2_keeps_the_other_species_state(oracle32, storage):
    """At mpdata_order 2 the fused path alternates between two buffers per
    species, so after one step the current state sits in the second one.
    Binding a new species then uploading only that species keeps the other
    species' device state: the next step continues from it, bit for bit as a
    handle that takes both steps without the new species (and, in FP32, as
    the oracle's two steps), and the new species takes one step from its
    uploaded values."""
    case = synthetic.config2(nx=40, ny=30, nz=8, steps=2, nspecies=3)
    case.options.kernel_variant = 1
    case.options.adv_options.mpdata_order = 2
    d, o = case.domain, case.options
    names = list(d.species)
    if storage != _lib.STORE_F32:
        d.storage[names[1]] = (storage, synthetic.pow2_scale(d.species[names[1]], storage))
    init = {n: a.copy() for n, a in d.species.items()}
    r = np.random.default_rng(41)
    tracer = r.uniform(0.5, 1.5, d.shape).astype(np.float32)
    tracer0 = tracer.copy()
    st = Advection(d, o)
    try:
        st.advect(case.dt)  # the state moves to the second buffers
        bind(st, 3, tracer)
        st._check(st.lib.icar_adv_upload(st.h, _lib.mask_species(3)))
        st.advect(case.dt)
        st.download()
    finally:
        st.close()
    got = {n: d.species[n].copy() for n in names}
    for n, a in init.items():  # the same two steps without the binding
        d.species[n][...] = a
    ref = Advection(d, o)
    try:
        ref.advect(case.dt)
        ref.advect(case.dt)
        ref.download()
    finally:
        ref.close()
    for n in names:
        np.testing.assert_array_equal(got[n].view(np.uint32), d.species[n].view(np.uint32),
                                      err_msg=n)
    fp32 = [n for n in names if n not in d.storage or d.storage[n][0] == _lib.STORE_F32]
    orc = oracle32.run(d, o, case.dt, 2, 0, 0, {n: init[n] for n in fp32})
    for n in fp32:
        np.testing.assert_array_equal(got[n].view(np.uint32), orc[n].view(np.uint32), err_msg=n)
    tr = oracle32.run(d, o, case.dt, 1, 0, 0, {"tracer": tracer0})["tracer"]
    np.testing.assert_array_equal(tracer.view(np.uint32), tr.view(np.uint32))
