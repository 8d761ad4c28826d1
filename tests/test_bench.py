"""The parts of bench.py that need no GPU: the one-GPU guard, the compulsory
bytes behind the roofline, the committed profile figures it reads, and the
shape of the CPU baseline it reports."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from hicar_amd import _lib, synthetic  # noqa: E402
from hicar_amd.domain import Domain  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_compulsory_bytes_per_cell_species():
    """4*(5N + 2N_u + 2N_v) + sum_s 2*b_s*N: 11.60 B per cell-species for the
    1024x1024x30 domain with 10 FP32 species (a shape-only domain: the
    bytes do not depend on the values)."""
    d = Domain(1024, 1024, 30, 1000.0)
    d.species.update({f"s{k}": None for k in range(10)})
    n = 1024 * 1024 * 30
    b = bench.step_bytes(d)
    assert b == 4 * (5 * n + 2 * 1025 * 30 * 1024 + 2 * 1024 * 30 * 1025) + 10 * 8 * n
    assert abs(b / (10 * n) - 11.6015625) < 1e-9
    # a 16-bit species moves 2 bytes each way
    d.storage["s0"] = (_lib.STORE_BF16, 0)
    assert bench.step_bytes(d) == b - 4 * n


def test_committed_profile_figures_are_read():
    """The roofline's traffic and VALU count come from profiles/: the
    corrective kernel's entries exist and are plausible for one launch on
    the 1024x1024x30 workload (more than its 5.03 GB of own bytes; billions
    of wave instructions)."""
    t = bench.read_traffic("k_corr", "c2_1024x1024x30")
    v = bench.read_valu("k_corr", "c2_1024x1024x30")
    assert t is not None and 5.0e9 < t < 2.0e10
    assert v is not None and 1.0e9 < v < 1.0e10


def test_cpu_baseline_object():
    """One image per thread, as ICAR runs one image per core: the threads
    used are the process's CPU share (at most 16), stated with the host's
    own counts."""
    case = synthetic.config2(nx=48, ny=40, nz=8, steps=1, nspecies=3)
    out = bench.cpu_baseline(case)
    assert set(out) >= {"value", "unit", "cores", "kind", "sample", "single_thread", "host"}
    assert out["kind"] == "port" and out["unit"] == "cell-species updates/s"
    share = len(os.sched_getaffinity(0))
    assert out["cores"] == min(16, share) and out["value"] > 0
    assert out["host"]["logical_cpus"] == os.cpu_count()
    assert out["single_thread"]["cores"] == 1
    json.dumps(out)


def test_cpu_baseline_images_cover_the_strip_and_step_like_the_whole():
    """The images' owned rows, stepped on their own with their 2-row halos,
    equal the whole strip's step away from the cuts' influence: an image
    is a subdomain, as ICAR's are (its ring, the halo, is not updated)."""
    import numpy as np
    from oracle.oracle import Oracle
    case = synthetic.config2(nx=40, ny=36, nz=6, steps=1, nspecies=2)
    d = case.domain
    orc = Oracle(np.float32)
    whole = orc.advect(d, case.options, case.dt)
    sub, sp = bench._image(d, 10, 24)
    part = orc.advect(sub, case.options, case.dt, species=sp)
    for n in whole:
        # rows 11..22 are updated inside the image from the same inputs; the
        # MPDATA stencil of the corrective pass reaches 2 rows, the limiter 1
        np.testing.assert_array_equal(part[n][3:-3], whole[n][13:21], err_msg=n)
