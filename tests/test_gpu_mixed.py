"""Reduced-precision species storage (BF16 / FP16 with a per-species power-of-
two scale) on the batched path.

Every pass computes in FP32 from the dequantised field and only the final
store of a step rounds (to nearest even) into the 16-bit format, so the GPU
result equals, bit for bit, the FP32 oracle stepped on the dequantised
values with the same rounding applied between steps (emulated here in
numpy).  Against the FP64 oracle the error of a 16-bit field is set by its
storage format (BF16 2^-9, FP16 2^-11 relative); qv and theta stay FP32.
"""
import numpy as np
import pytest

from hicar_amd import _lib, synthetic
from hicar_amd.advection import MpdataAdvection

from test_gpu_parity import fresh_species

pytestmark = pytest.mark.gpu


def quantize(a: np.ndarray, storage: int, e: int) -> np.ndarray:
    """value -> 16-bit storage (round to nearest even) -> value, as the
    device stores it."""
    x = (np.asarray(a, np.float32) * np.float32(2.0 ** -e)).astype(np.float32)
    if storage == _lib.STORE_F16:
        y = x.astype(np.float16).astype(np.float32)
    else:
        u = x.view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
        y = (u.astype(np.uint32) << 16).view(np.float32)
    return (y * np.float32(2.0 ** e)).astype(np.float32)


def emulate(oracle, case, steps, dtype=np.float32):
    d = case.domain
    cur = {}
    for n, a in d.species.items():
        st, e = d.storage.get(n, (_lib.STORE_F32, 0))
        cur[n] = quantize(a, st, e) if st != _lib.STORE_F32 else a.astype(np.float32).copy()
    for _ in range(steps):
        cur = oracle.run(d, case.options, case.dt, 1, case.bc_x, case.bc_y, cur)
        for n in cur:
            st, e = d.storage.get(n, (_lib.STORE_F32, 0))
            if st != _lib.STORE_F32:
                cur[n] = quantize(cur[n], st, e)
    return cur


def run_gpu(case, steps, budget=False):
    st = MpdataAdvection(case.domain, case.options)
    try:
        st.run(case.dt, steps, case.bc_x, case.bc_y)
        st.download()
        out = {n: case.domain.species[n].copy() for n in st.names}
        return (out, dict(zip(st.names, st.mass_budget()))) if budget else out
    finally:
        st.close()


def oracle_run_threaded(oracle, case, steps, species):
    """The oracle on each species in its own thread (species are independent
    inside advect; the C oracle keeps no state and ctypes releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    d = case.domain

    def one(name):
        return name, oracle.run(d, case.options, case.dt, steps, case.bc_x, case.bc_y,
                                {name: species[name]})[name]

    with ThreadPoolExecutor(max_workers=min(16, len(species))) as ex:
        return dict(ex.map(one, list(species)))


@pytest.mark.parametrize("storage", [_lib.STORE_BF16, _lib.STORE_F16])
def test_mixed_storage_matches_emulated_oracle(oracle32, storage):
    case = synthetic.config4(storage=storage, nx=96, ny=40, nz=20, steps=15)
    ref = emulate(oracle32, case, case.steps)
    got = run_gpu(case, case.steps)
    for n in ref:
        np.testing.assert_array_equal(got[n], ref[n], err_msg=n)
    # the 16-bit fields really are stored in 16 bits
    assert case.domain.storage["qc"][0] == storage


def test_mixed_c4_error_against_fp64(oracle64, capsys):
    """The 256x64x20, 12-species domain with BF16 (then FP16) hydrometeors,
    200 steps, against the FP64 oracle: per-field normalised max and mean
    error, and the drift of the FP64 mass budget of each species from the
    FP64 oracle's.  A 16-bit field is re-rounded every step, so its error
    grows like a random walk of the unit roundoff u (BF16 2^-9, FP16 2^-11):
    bounded by 4*sqrt(steps)*u; FP32 fields by 1e-4.  Round-to-nearest-even
    is unbiased, so the mass drift stays far below u: bounded by 1e-4 for
    every field."""
    steps = 200
    base = synthetic.config4(storage=_lib.STORE_BF16, steps=steps)
    d = base.domain
    ref = oracle_run_threaded(oracle64, base, steps, fresh_species(base))
    f64 = lambda a: np.asarray(a, np.float64)
    ref_mass = {n: oracle64.mass(r, f64(d.jacobian), f64(d.density), f64(d.advection_dz))
                for n, r in ref.items()}
    report = []
    for storage, unit in ((_lib.STORE_BF16, 2.0 ** -9), (_lib.STORE_F16, 2.0 ** -11)):
        floor = 4.0 * np.sqrt(steps) * unit
        case = synthetic.config4(storage=storage, steps=steps)
        got, mass = run_gpu(case, steps, budget=True)
        for n, r in ref.items():
            g = got[n].astype(np.float64)
            emax = np.abs(g - r).max() / np.abs(r).max()
            emean = np.abs(g - r).mean() / np.abs(r).mean()
            dm = abs(mass[n] - ref_mass[n]) / abs(ref_mass[n])
            st = case.domain.storage.get(n, (_lib.STORE_F32, 0))[0]
            report.append((storage, n, st, emax, emean, dm))
            bound = 1e-4 if st == _lib.STORE_F32 else floor
            assert emax <= bound, (storage, n, emax)
            assert dm <= 1e-4, (storage, n, dm)
    with capsys.disabled():
        for storage, n, st, emax, emean, dm in report:
            fmt = {0: "f32", 1: "bf16", 2: "f16"}[st]
            run = {1: "bf16-run", 2: "f16-run"}[storage]
            print(f"\n[precision] {run} {n:6s} {fmt:5s} max {emax:.2e} mean {emean:.2e} "
                  f"mass drift {dm:.2e}", end="")
        print()
