"""Reduced-precision species storage (BF16 / FP16 with a per-species power-of-
two scale, rounded to nearest even or stochastically) on the batched path.

Every pass computes in FP32 from the dequantised field and only the final
store of a step rounds into the 16-bit format, so the GPU result equals, bit
for bit, the FP32 oracle stepped on the dequantised values with the same
rounding applied between steps (emulated here in numpy, including the
counter-based random bits of the stochastic kinds).  Against the FP64 oracle
the error of a 16-bit field is set by its storage format (BF16 2^-9, FP16
2^-11 relative); qv and theta stay FP32.  Round-to-nearest storage loses
updates smaller than half an ulp and so drifts in mass; stochastic rounding
keeps the mass of a closed box within its random-walk bound.
"""
import numpy as np
import pytest

from hicar_amd import _lib, synthetic
from hicar_amd.advection import MpdataAdvection

from test_gpu_parity import fresh_species

pytestmark = pytest.mark.gpu

SR_KINDS = (_lib.STORE_BF16_SR, _lib.STORE_F16_SR)
KINDS16 = (_lib.STORE_BF16, _lib.STORE_F16, _lib.STORE_BF16_SR, _lib.STORE_F16_SR)
_M32 = np.uint64(0xFFFFFFFF)


def is_bf16(storage):
    return storage in (_lib.STORE_BF16, _lib.STORE_BF16_SR)


def quantize(a: np.ndarray, storage: int, e: int) -> np.ndarray:
    """value -> 16-bit storage (round to nearest even) -> value, as the
    device stores it on upload and for the plain 16-bit kinds."""
    x = (np.asarray(a, np.float32) * np.float32(2.0 ** -e)).astype(np.float32)
    if not is_bf16(storage):
        y = x.astype(np.float16).astype(np.float32)
    else:
        u = x.view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
        y = (u.astype(np.uint32) << 16).view(np.float32)
    return (y * np.float32(2.0 ** e)).astype(np.float32)


def sr_bits(n: int, s: int, t: int) -> np.ndarray:
    """The device's 16 random bits of cells 0..n-1, species s, step t: the
    murmur3 32-bit finaliser of (c*0x9E3779B1) ^ ((s+1)*0x85EBCA77) ^
    (t*0xC2B2AE3D)."""
    c = np.arange(n, dtype=np.uint64)
    x = ((c * np.uint64(0x9E3779B1)) & _M32) ^ np.uint64(((s + 1) * 0x85EBCA77) & 0xFFFFFFFF) \
        ^ np.uint64((t * 0xC2B2AE3D) & 0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x85EBCA6B)) & _M32
    x ^= x >> np.uint64(13)
    x = (x * np.uint64(0xC2B2AE35)) & _M32
    x ^= x >> np.uint64(16)
    return (x & np.uint64(0xFFFF)).astype(np.uint32)


def quantize_sr(a: np.ndarray, storage: int, e: int, s: int, t: int) -> np.ndarray:
    """value -> 16-bit storage with stochastic rounding -> value.  The flat
    C index of a (ny, nz, nx) array is the cell index i + nx*(k + nz*j)."""
    x = (np.asarray(a, np.float32) * np.float32(2.0 ** -e)).astype(np.float32).reshape(-1)
    r = sr_bits(x.size, s, t)
    if is_bf16(storage):
        u = (x.view(np.uint32).astype(np.uint64) + r) >> np.uint64(16)
        y = (u.astype(np.uint32) << 16).view(np.float32)
    else:
        ax = np.abs(x)
        u = ((x.view(np.uint32).astype(np.uint64) + (r & np.uint32(0x1FFF))) & _M32)
        yn = ((u.astype(np.uint32) >> 13) << 13).view(np.float32)  # truncate to FP16 normal
        a24 = ax * np.float32(2.0 ** 24)
        k = np.floor(a24)
        up = (r.astype(np.float32) * np.float32(2.0 ** -16)) < (a24 - k)
        ys = np.copysign((k + up) * np.float32(2.0 ** -24), x).astype(np.float32)
        y = np.where(ax >= np.float32(2.0 ** -14), yn, ys).astype(np.float32)
    return (y.reshape(np.shape(a)) * np.float32(2.0 ** e)).astype(np.float32)


def store(a, storage, e, s, t):
    if storage in SR_KINDS:
        return quantize_sr(a, storage, e, s, t)
    return quantize(a, storage, e)


def emulate(oracle, case, steps):
    d = case.domain
    cur = {}
    for n, a in d.species.items():
        st, e = d.storage.get(n, (_lib.STORE_F32, 0))
        cur[n] = quantize(a, st, e) if st != _lib.STORE_F32 else a.astype(np.float32).copy()
    for t in range(steps):
        # the device rounds in the last pass and then fills the ring from
        # the stored (rounded) values
        cur = oracle.advect(d, case.options, case.dt, cur)
        for s, n in enumerate(cur):
            st, e = d.storage.get(n, (_lib.STORE_F32, 0))
            if st != _lib.STORE_F32:
                cur[n] = store(cur[n], st, e, s, t)
            oracle.apply_bc(cur[n], case.bc_x, case.bc_y)
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


_C4_FP64 = {}


def c4_fp64_reference(oracle64, steps):
    """The FP64 oracle's run of the 256x64x20, 12-species domain (computed
    once per session: two tests compare with it)."""
    if steps not in _C4_FP64:
        base = synthetic.config4(storage=_lib.STORE_BF16, steps=steps)
        _C4_FP64[steps] = oracle_run_threaded(oracle64, base, steps, fresh_species(base))
    return _C4_FP64[steps]


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


@pytest.mark.parametrize("storage", KINDS16)
def test_mixed_storage_matches_emulated_oracle(oracle32, storage):
    case = synthetic.config4(storage=storage, nx=96, ny=40, nz=20, steps=15)
    ref = emulate(oracle32, case, case.steps)
    got = run_gpu(case, case.steps)
    for n in ref:
        np.testing.assert_array_equal(got[n], ref[n], err_msg=n)
    # the 16-bit fields really are stored in 16 bits
    assert case.domain.storage["qc"][0] == storage


@pytest.mark.parametrize("storage", KINDS16)
@pytest.mark.parametrize("fct", [True, False])
def test_mixed_storage_order3_matches_emulated_oracle(oracle32, storage, fct):
    """Order 3: the first corrective pass reads the 16-bit q0 and writes an
    FP32 intermediate, the second reads both and writes the 16-bit result.
    Each 16-bit class of the kernels (the BF16 kinds beside FP32, any kind
    with FP16) against the FP32 oracle stepped on the 16-bit values, bit for
    bit, with and without FCT (without it species can go negative)."""
    case = synthetic.config4(storage=storage, nx=70, ny=36, nz=14, steps=6)
    case.options.adv_options.mpdata_order = 3
    case.options.adv_options.MPDATA_FCT = fct
    ref = emulate(oracle32, case, case.steps)
    got = run_gpu(case, case.steps)
    for n in ref:
        np.testing.assert_array_equal(got[n], ref[n], err_msg=n)


@pytest.mark.parametrize("order", [2, 3])
def test_every_kind_in_one_batch_matches_emulated_oracle(oracle32, order):
    """One domain whose species use every storage kind at once (FP32, BF16,
    BF16 stochastic, FP16, FP16 stochastic): the kernels for any kind, where
    each block branches on its species' kind, against the FP32 oracle
    stepped on the stored values, bit for bit."""
    case = synthetic.config4(storage=_lib.STORE_BF16_SR, nx=80, ny=36, nz=16, steps=6)
    case.options.adv_options.mpdata_order = order
    d = case.domain
    kinds = (_lib.STORE_BF16, _lib.STORE_F16, _lib.STORE_F16_SR, _lib.STORE_BF16_SR, None)
    for idx, n in enumerate([n for n in d.species if n not in ("qv", "theta")]):
        kind = kinds[idx % len(kinds)]
        if kind is None:
            d.storage.pop(n, None)  # FP32 between 16-bit species
        else:
            d.storage[n] = (kind, synthetic.pow2_scale(d.species[n], kind))
    assert {k for k, _ in d.storage.values()} == set(KINDS16)
    ref = emulate(oracle32, case, case.steps)
    got = run_gpu(case, case.steps)
    for n in ref:
        np.testing.assert_array_equal(got[n], ref[n], err_msg=n)


def test_ensemble_of_storage_classes_matches_each_member(oracle32):
    """An ensemble whose members hold an FP32-only, a BF16 and an FP16
    species set runs the kernels of the widest class for all of them; each
    member equals its own run (and so, through the tests above, the
    oracle), bit for bit."""
    from hicar_amd.advection import Ensemble
    cases = [synthetic.config4(storage=st, nx=nx, ny=30, nz=12, steps=4)
             for st, nx in ((_lib.STORE_BF16_SR, 70), (_lib.STORE_F16, 54), (_lib.STORE_BF16, 40))]
    cases[2].domain.storage.clear()  # every species FP32
    want = []
    for c in cases:
        init = {n: a.copy() for n, a in c.domain.species.items()}
        want.append(run_gpu(c, c.steps))
        for n, a in init.items():
            c.domain.species[n][...] = a
    members = [MpdataAdvection(c.domain, c.options) for c in cases]
    ens = Ensemble(members)
    try:
        ens.run([c.dt for c in cases], cases[0].steps, 1, 2, graph=True)
        ens.sync()
        for m, c, w in zip(members, cases, want):
            m.download()
            for n in w:
                np.testing.assert_array_equal(c.domain.species[n], w[n], err_msg=n)
    finally:
        ens.close()
        for m in members:
            m.close()


def test_mixed_c4_error_against_fp64(oracle64, capsys):
    """The 256x64x20, 12-species domain with its hydrometeors in each 16-bit
    kind, 200 steps, against the FP64 oracle: per-field normalised max and
    mean error, and the difference of the FP64 mass budget from the FP64
    oracle's.  A 16-bit field is re-rounded every step, so its error grows
    like a random walk of the unit roundoff u (BF16 2^-9, FP16 2^-11):
    bounded by 4*sqrt(steps)*u; FP32 fields by 1e-4 and their budget by
    1e-6.  The domain is open (inflow, outflow, lid), so the budget of a
    16-bit field also moves with its trajectory's boundary fluxes: reported,
    while conservation itself is asserted in the closed box below."""
    steps = 200
    base = synthetic.config4(storage=_lib.STORE_BF16, steps=steps)
    d = base.domain
    ref = c4_fp64_reference(oracle64, steps)
    f64 = lambda a: np.asarray(a, np.float64)
    ref_mass = {n: oracle64.mass(r, f64(d.jacobian), f64(d.density), f64(d.advection_dz))
                for n, r in ref.items()}
    report = []
    for storage in KINDS16:
        unit = 2.0 ** -9 if is_bf16(storage) else 2.0 ** -11
        floor = 4.0 * np.sqrt(steps) * unit
        case = synthetic.config4(storage=storage, steps=steps)
        got, mass = run_gpu(case, steps, budget=True)
        for n, r in ref.items():
            g = got[n].astype(np.float64)
            emax = np.abs(g - r).max() / np.abs(r).max()
            emean = np.abs(g - r).mean() / np.abs(r).mean()
            dm = (mass[n] - ref_mass[n]) / abs(ref_mass[n])
            st = case.domain.storage.get(n, (_lib.STORE_F32, 0))[0]
            report.append((storage, n, st, emax, emean, dm))
            if st == _lib.STORE_F32:
                assert emax <= 1e-4, (storage, n, emax)
                assert abs(dm) <= 1e-6, (storage, n, dm)
            else:
                assert emax <= floor, (storage, n, emax)
    with capsys.disabled():
        for storage, n, st, emax, emean, dm in report:
            print(f"\n[precision] {_lib.STORE_NAMES[storage]}-run {n:6s} {_lib.STORE_NAMES[st]:7s} "
                  f"max {emax:.2e} mean {emean:.2e} budget {dm:+.2e}", end="")
        print()


def closed_box_case(storage, nx=70, ny=50, nz=20, steps=200):
    """The 10-species workload's fields in a closed box: a horizontal
    streamfunction flow with no flux through the ring, w = 0, rho = 1."""
    case = synthetic.config2(nx=nx, ny=ny, nz=nz, steps=steps)
    d = case.domain
    d.w[...] = 0
    d.u[...] = 0
    d.v[...] = 0
    rng = np.random.default_rng(3)
    psi = np.zeros((d.ny + 1, d.nz, d.nx + 1))
    psi[2:-2, :, 2:-2] = rng.uniform(-200, 200, psi[2:-2, :, 2:-2].shape)
    d.u[:, :, 1:-1] = (psi[1:, :, 1:-1] - psi[:-1, :, 1:-1]) / d.dx * 10
    d.v[1:-1] = -(psi[1:-1, :, 1:] - psi[1:-1, :, :-1]) / d.dx * 10
    d.density[...] = 1.0
    d.jacobian_u[...] = 1.0
    d.jacobian_v[...] = 1.0
    case.options.kernel_variant = 1
    for n, a in d.species.items():
        if n not in ("qv", "theta"):
            d.storage[n] = (storage, synthetic.pow2_scale(a, storage))
    return case


def ulp16(q, storage, e):
    """Spacing of the 16-bit grid around each value (2^e scaled)."""
    x = np.abs(np.asarray(q, np.float64)) * 2.0 ** -e
    p = 8 if is_bf16(storage) else 11
    ex = np.floor(np.log2(np.maximum(x, 1e-300)))
    if not is_bf16(storage):
        ex = np.maximum(ex, -14)
    return np.where(x > 0, 2.0 ** (ex - p + 1), 0.0) * 2.0 ** e


@pytest.mark.parametrize("storage", KINDS16)
def test_mixed_closed_box_mass(storage, capsys):
    """Closed box, 200 steps, order 3 with FCT.  With stochastic rounding the
    FP64 budget of every 16-bit field moves only by the random walk of the
    per-step rounding: within 6 sigma, sigma^2 = steps * sum((ulp*G*dz)^2)/4
    from the field's ulps.  Round-to-nearest drifts (reported)."""
    steps = 200
    case = closed_box_case(storage, steps=steps)
    d = case.domain
    st = MpdataAdvection(d, case.options)
    try:
        m0 = st.mass_budget()
        st.run(case.dt, steps)
        m1 = st.mass_budget()
        st.download()
    finally:
        st.close()
    w = (np.asarray(d.jacobian, np.float64) * np.asarray(d.density, np.float64)
         * np.asarray(d.advection_dz, np.float64))[1:-1, :, 1:-1]
    lines = []
    for s, n in enumerate(d.species):
        kind, e = d.storage.get(n, (_lib.STORE_F32, 0))
        rel = (m1[s] - m0[s]) / abs(m0[s])
        if kind == _lib.STORE_F32:
            assert abs(rel) < 2e-6, (n, rel)
            continue
        u = ulp16(d.species[n], kind, e)[1:-1, :, 1:-1]
        sigma = np.sqrt(steps * np.sum((u * w) ** 2) / 4.0)
        lines.append(f"[closed box] {_lib.STORE_NAMES[kind]:7s} {n:4s} budget {rel:+.2e} "
                     f"(6 sigma {6 * sigma / abs(m0[s]):.1e})")
        if kind in SR_KINDS:
            assert abs(m1[s] - m0[s]) <= 6 * sigma, (n, rel, sigma / abs(m0[s]))
    with capsys.disabled():
        print("\n" + "\n".join(lines))


@pytest.mark.parametrize("workload", ["c2", "c5"])
def test_mixed_crops_error_against_fp64(oracle64, workload, capsys):
    """Per-field error against the FP64 oracle on crops of the two large
    workloads, 20 steps, with every species in FP32 and with the
    hydrometeors stored in 16 bits as the benchmark stores them:
      c2  130 x 70 x 30 of the 10-species domain, order 3 + FCT; BF16 with
          stochastic rounding for all but qv and theta;
      c5  192 x 192 x 50 of the 4096 x 4096 x 50 domain (dx 250 m, jet and
          ridge), 12 species, order 2 + FCT; ten in BF16, stochastic
          rounding.
    FP32 fields meet 1e-4 relative to max |q|, except theta on c2: at
    order 3 with FCT its top levels diverge from FP64 in the reference's own
    FP32 arithmetic (test_precision_spread.py), so there the bound holds
    below the top 10 levels.  A 16-bit field stays within 4*sqrt(steps)*2^-9.
    The lines printed are the report."""
    steps = 20
    if workload == "c2":
        def make():
            c = synthetic.config2(nx=130, ny=70, nz=30, steps=steps)
            c.options.kernel_variant = 1
            return c
    else:
        make = lambda: synthetic.config5(nx=192, ny=192, nz=50, steps=steps)
    base = make()
    d = base.domain
    base.domain.storage.clear()
    ref = oracle_run_threaded(oracle64, base, steps, fresh_species(base))
    f64 = lambda a: np.asarray(a, np.float64)
    ref_mass = {n: oracle64.mass(r, f64(d.jacobian), f64(d.density), f64(d.advection_dz))
                for n, r in ref.items()}
    report = []
    for mixed in (False, True):
        case = make()
        case.domain.storage.clear()
        if mixed:
            kind = _lib.STORE_BF16_SR
            for n, a in case.domain.species.items():
                if n not in ("qv", "theta"):
                    case.domain.storage[n] = (kind, synthetic.pow2_scale(a, kind))
        got, mass = run_gpu(case, steps, budget=True)
        nz = case.domain.nz
        for n, r in ref.items():
            g = got[n].astype(np.float64)
            emax = np.abs(g - r).max() / np.abs(r).max()
            emean = np.abs(g - r).mean() / np.abs(r).mean()
            dm = (mass[n] - ref_mass[n]) / abs(ref_mass[n])
            st = case.domain.storage.get(n, (_lib.STORE_F32, 0))[0]
            report.append((mixed, n, st, emax, emean, dm))
            if st == _lib.STORE_F32:
                if workload == "c2" and n == "theta":
                    below = np.abs(g[:, :nz - 10] - r[:, :nz - 10]).max() / np.abs(r).max()
                    assert below <= 1e-4, (mixed, n, below)
                else:
                    assert emax <= 1e-4, (mixed, n, emax)
            else:
                assert emax <= 4.0 * np.sqrt(steps) * 2.0 ** -9, (mixed, n, emax)
    with capsys.disabled():
        for mixed, n, st, emax, emean, dm in report:
            print(f"\n[precision {workload}] {'mixed' if mixed else 'fp32 '} {n:6s} "
                  f"{_lib.STORE_NAMES[st]:7s} max {emax:.2e} mean {emean:.2e} budget {dm:+.2e}",
                  end="")
        print()


def test_compensated_divergence_against_fp64(oracle64, capsys):
    """The opt-in compensated flux divergence (icar_adv_set_compensation)
    against the FP64 oracle on the 256x64x20, 12-species domain, 200 steps,
    with every species in FP32 and with the hydrometeors in BF16 (stochastic
    rounding), beside the same runs in the reference's operation order.
    The lines printed are the measurement: per field, the max error
    relative to max |q| with and without compensation.  Asserted: the
    compensated runs stay within the same bounds as the plain ones (FP32
    fields 1e-4, 16-bit fields 4*sqrt(steps)*2^-9), and the FP32 compensated
    run really differs from the plain one (it is not the reference's
    rounding)."""
    steps = 200
    ref = c4_fp64_reference(oracle64, steps)
    rows = {}
    runs = {}
    for kind in ("f32", "bf16-sr"):
        for comp in (False, True):
            case = synthetic.config4(storage=_lib.STORE_BF16_SR, steps=steps)
            if kind == "f32":
                case.domain.storage.clear()
            st = MpdataAdvection(case.domain, case.options)
            try:
                st.set_compensation(comp)
                st.run(case.dt, steps, case.bc_x, case.bc_y)
                st.download()
            finally:
                st.close()
            got = {n: case.domain.species[n].copy() for n in case.domain.species}
            runs[(kind, comp)] = got
            for n, r in ref.items():
                e = np.abs(got[n].astype(np.float64) - r).max() / np.abs(r).max()
                rows.setdefault((kind, n), {})[comp] = e
                st16 = case.domain.storage.get(n, (_lib.STORE_F32, 0))[0]
                bound = 1e-4 if st16 == _lib.STORE_F32 else 4.0 * np.sqrt(steps) * 2.0 ** -9
                assert e <= bound, (kind, comp, n, e)
    assert any(not np.array_equal(runs[("f32", True)][n], runs[("f32", False)][n])
               for n in ref)
    with capsys.disabled():
        for (kind, n), e in rows.items():
            print(f"\n[compensation] {kind:7s} {n:6s} max err plain {e[False]:.3e} "
                  f"compensated {e[True]:.3e}", end="")
        print()


@pytest.mark.parametrize("cfg,ksplit", [(0, 1), (3, 3), (5, 1), (7, 2), (1, 4)])
def test_mixed_storage_every_launch_configuration(oracle32, monkeypatch, cfg, ksplit):
    """The BF16 class of the corrective kernel under each register budget,
    prefetch choice and tile height (60x16, 60x32), with the columns split
    into chunks (whose warm-up levels read the 16-bit q0), against the
    emulated oracle bit for bit; order 3 so that both corrective passes run."""
    monkeypatch.setenv("ICAR_ADV_CORR_CFG", str(cfg))
    monkeypatch.setenv("ICAR_ADV_KSPLIT", str(ksplit))
    case = synthetic.config4(storage=_lib.STORE_BF16_SR, nx=130, ny=70, nz=20, steps=4)
    case.options.adv_options.mpdata_order = 3
    ref = emulate(oracle32, case, case.steps)
    st = MpdataAdvection(case.domain, case.options)
    try:
        st.run(case.dt, case.steps, case.bc_x, case.bc_y)
        st.download()
        assert tuple(st.launch_config()) == (cfg, ksplit)
        got = {n: case.domain.species[n].copy() for n in st.names}
    finally:
        st.close()
    for n in ref:
        np.testing.assert_array_equal(got[n], ref[n], err_msg=n)


@pytest.mark.parametrize("storage", [_lib.STORE_BF16_SR, _lib.STORE_F16_SR])
def test_graph_replays_advance_the_stochastic_rounding(oracle32, storage):
    """A recorded step loop with stochastically rounded species, replayed
    three times with downloads between the replays: the step counter that
    seeds the rounding advances inside the graph, so the result equals the
    emulated oracle over all the steps, bit for bit."""
    case = synthetic.config4(storage=storage, nx=70, ny=40, nz=12, steps=3)
    ref = emulate(oracle32, case, 3 * case.steps)
    st = MpdataAdvection(case.domain, case.options)
    try:
        st.set_graph(True)
        for _ in range(3):
            st.run(case.dt, case.steps, case.bc_x, case.bc_y)
            st.download()  # host work between the replays
        got = {n: case.domain.species[n].copy() for n in st.names}
    finally:
        st.close()
    for n in ref:
        np.testing.assert_array_equal(got[n], ref[n], err_msg=n)
