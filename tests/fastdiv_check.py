"""Parity of the fused kernels with the opt-in three-operation quotient
(ICAR_ADV_FAST_DIV=1, read once per process): run as a script in a process
of its own by tests/test_gpu_division.py.  Random domains (orders 2-4, FCT
on and off, edge tiles), species spanning 1e-30..1e25 and of both signs (the
guard's fallback to the IEEE division), the 130x70x30 crop of the 10-species
domain for 10 steps, and 16-bit storage against the emulated oracle: every
case bit for bit.  Prints one line per case; exit status 1 on a mismatch."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

from hicar_amd import _lib, synthetic  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402
from test_gpu_mixed import emulate, run_gpu  # noqa: E402
from test_gpu_parity import fresh_species, gpu_run, random_domain  # noqa: E402


def same(got, ref):
    return all(np.array_equal(got[n], ref[n]) for n in ref)


def main() -> int:
    assert os.environ.get("ICAR_ADV_FAST_DIV") == "1"
    orc = Oracle(np.float32)
    bad = 0
    cases = []
    for order, fct in ((2, True), (3, True), (4, True), (3, False)):
        cases.append((f"random order {order} fct {fct}",
                      random_domain(131, 37, 9, 500 + order, order, fct, True, nspec=3)))
    ext = random_domain(131, 37, 9, 77, 3, False, True, nspec=4)
    names = list(ext.domain.species)
    r = np.random.default_rng(78)
    ext.domain.species[names[0]][...] *= np.float32(1e-30)
    ext.domain.species[names[1]][...] *= np.float32(1e25)
    ext.domain.species[names[2]][...] = (10.0 ** r.uniform(-38, 20, ext.domain.shape)).astype(
        np.float32) * (r.random(ext.domain.shape) > 0.3)
    ext.domain.species[names[3]][...] -= np.float32(0.5)
    cases.append(("extreme values, both signs, no FCT", ext))
    cases.append(("130x70x30 crop, 10 steps", synthetic.config2(nx=130, ny=70, nz=30, steps=10)))
    for name, case in cases:
        ref = orc.run(case.domain, case.options, case.dt, case.steps, case.bc_x, case.bc_y,
                      fresh_species(case))
        ok = same(gpu_run(case, variant=1), ref)
        bad += not ok
        print(f"fast division: {name}: {'bit-identical' if ok else 'DIFFERS'}", flush=True)
    case = synthetic.config4(storage=_lib.STORE_BF16_SR, nx=96, ny=40, nz=20, steps=10)
    ref = emulate(orc, case, case.steps)  # first: the GPU run downloads into the domain
    ok = same(run_gpu(case, case.steps), ref)
    bad += not ok
    print(f"fast division: bf16-sr storage: {'bit-identical' if ok else 'DIFFERS'}", flush=True)
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
