"""icar_adv_mass_budget_noalloc, the entry point for callers that cannot
allocate: its FP64 budget of one FP32 species on the caller's device arrays
equals icar_adv_mass_budget's for the same values bit for bit, on two calls
in a row (the second on a smaller domain, so the fixed partial-sum array
holds the first call's values beyond the second's blocks), eagerly and
replayed from a captured graph; bad arguments return errors.  The caller's
device memory, stream and graph come from the HIP runtime the library
itself uses (plumbing for the test)."""
import ctypes
import os

import numpy as np
import pytest

from hicar_amd import _lib, synthetic
from hicar_amd.advection import Advection

pytestmark = pytest.mark.gpu

vp = ctypes.c_void_p
H2D, D2H = 1, 2


class Hip:
    """The few HIP runtime calls the test needs, through ctypes."""

    def __init__(self):
        _lib.load()  # the library's own runtime is then already in the process
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        self.h = ctypes.CDLL(path)
        self.bufs = []

    def check(self, rc, what):
        assert rc == 0, f"{what}: hipError {rc}"

    def to_device(self, a: np.ndarray) -> vp:
        a = np.ascontiguousarray(a)
        p = vp()
        self.check(self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(a.nbytes)), "hipMalloc")
        self.bufs.append(p)
        self.check(self.h.hipMemcpy(p, a.ctypes.data_as(vp), ctypes.c_size_t(a.nbytes), H2D),
                   "hipMemcpy")
        return p

    def from_device(self, p, n, dtype) -> np.ndarray:
        out = np.zeros(n, dtype)
        self.check(self.h.hipMemcpy(out.ctypes.data_as(vp), p, ctypes.c_size_t(out.nbytes), D2H),
                   "hipMemcpy")
        return out

    def sync(self):
        self.check(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def close(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs.clear()


def handle_budget(case):
    st = Advection(case.domain, case.options)
    try:
        return st.mass_budget()
    finally:
        st.close()


def device_inputs(hip, case):
    d = case.domain
    dev = {n: hip.to_device(np.asarray(getattr(d, n), np.float32))
           for n in ("jacobian", "density", "advection_dz")}
    dev["q"] = [hip.to_device(a) for a in d.species.values()]
    return dev


def call(lib, case, dev, s, out, stream):
    d = case.domain
    return lib.icar_adv_mass_budget_noalloc(d.nx, d.ny, d.nz, dev["q"][s], dev["jacobian"],
                                            dev["density"], dev["advection_dz"], out, stream)


def test_noalloc_budget_equals_the_handle_budget_on_two_calls():
    lib = _lib.load()
    hip = Hip()
    try:
        big = synthetic.config2(nx=300, ny=260, nz=20, steps=1, nspecies=3)
        small = synthetic.config2(nx=40, ny=30, nz=8, steps=1, nspecies=2)
        for case in (big, small):  # more partial sums, then fewer
            ref = handle_budget(case)
            dev = device_inputs(hip, case)
            S = len(case.domain.species)
            out = hip.to_device(np.zeros(S, np.float64))
            for s in range(S):
                rc = call(lib, case, dev, s, vp(out.value + 8 * s), None)
                assert rc == 0, lib.icar_adv_last_error(None)
            hip.sync()
            got = hip.from_device(out, S, np.float64)
            np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))
    finally:
        hip.close()


def test_noalloc_budget_replayed_from_a_graph():
    """Captured once on a stream of the caller's, replayed three times with
    new values copied in and synchronisations between the replays; every
    replay equals the handle's budget of the values the arrays then hold."""
    lib = _lib.load()
    hip = Hip()
    h = hip.h
    stream, graph, gexec = vp(), vp(), vp()
    try:
        case = synthetic.config2(nx=64, ny=48, nz=10, steps=1, nspecies=1)
        dev = device_inputs(hip, case)
        out = hip.to_device(np.zeros(1, np.float64))
        hip.check(h.hipStreamCreate(ctypes.byref(stream)), "hipStreamCreate")
        # the first call on a device looks up the fixed array's address:
        # outside the capture
        assert call(lib, case, dev, 0, out, stream) == 0
        hip.check(h.hipStreamSynchronize(stream), "hipStreamSynchronize")
        hip.check(h.hipStreamBeginCapture(stream, 1), "hipStreamBeginCapture")
        rc = call(lib, case, dev, 0, out, stream)
        hip.check(h.hipStreamEndCapture(stream, ctypes.byref(graph)), "hipStreamEndCapture")
        assert rc == 0, lib.icar_adv_last_error(None)
        hip.check(h.hipGraphInstantiate(ctypes.byref(gexec), graph, None, None, ctypes.c_size_t(0)),
                  "hipGraphInstantiate")
        name = list(case.domain.species)[0]
        q0 = case.domain.species[name].copy()
        nbytes = ctypes.c_size_t(q0.nbytes)
        for scale in (1.0, 0.5, 3.0):
            host = (q0 * np.float32(scale)).astype(np.float32)
            hip.check(h.hipMemcpy(dev["q"][0], host.ctypes.data_as(vp), nbytes, H2D), "hipMemcpy")
            hip.check(h.hipGraphLaunch(gexec, stream), "hipGraphLaunch")
            hip.check(h.hipStreamSynchronize(stream), "hipStreamSynchronize")
            case.domain.species[name][...] = host
            ref = handle_budget(case)
            got = hip.from_device(out, 1, np.float64)
            assert got.view(np.uint64)[0] == ref.view(np.uint64)[0], scale
    finally:
        if gexec.value:
            h.hipGraphExecDestroy(gexec)
        if graph.value:
            h.hipGraphDestroy(graph)
        if stream.value:
            h.hipStreamDestroy(stream)
        hip.close()


def test_noalloc_budget_rejects_bad_arguments():
    lib = _lib.load()
    hip = Hip()
    try:
        case = synthetic.config2(nx=20, ny=16, nz=6, steps=1, nspecies=1)
        dev = device_inputs(hip, case)
        out = hip.to_device(np.zeros(1, np.float64))
        rc = lib.icar_adv_mass_budget_noalloc(2, 16, 6, dev["q"][0], dev["jacobian"],
                                              dev["density"], dev["advection_dz"], out, None)
        assert rc == _lib.EINVAL and b">= 3" in lib.icar_adv_last_error(None)
        rc = lib.icar_adv_mass_budget_noalloc(20, 16, 6, None, dev["jacobian"], dev["density"],
                                              dev["advection_dz"], out, None)
        assert rc == _lib.EINVAL and b"NULL" in lib.icar_adv_last_error(None)
    finally:
        hip.close()
