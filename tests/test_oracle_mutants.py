"""Mutation check of the oracle's pinning tests (CPU only).

Each mutant below is a plausible misreading of the reference planted in a
scratch copy of the oracle's C sources; the copy is built and the pinning
tests are run against it.  Every mutant must fail at least one of them, and
the unmutated oracle passes them all (the rest of the suite runs it).

  fct_start, fct_end   the FCT core limits the x/y boundary cell as it does
                       a w column's end cell (adv_mpdata_FCT_core.f90:79-83,
                       :95-99 set fin = fout = 0 there instead)
  wrf_x_both, _y_both  the WRF limiter takes both neighbours of a cell with
                       inflow through both faces (adv_fluxcorr.f90:99-113 is
                       `where ... else where`: the west/south one only)
  wrf_w_above          the level below enters the bounds on the sign of the
                       face above (:115-117 use the face below)
  wrf_w_below          the level above enters the bounds on the sign of the
                       face below (:118-120 use the face above)
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MUTANTS = {
    "fct_start": (
        "mpdata_oracle.c",
        "        } else {\n            fin_i = 0;\n            fout_i = 0;\n        }",
        "        } else {\n            fin_i = L(0.) - rmin2(L(0.), f[i]);\n"
        "            fout_i = rmax2(L(0.), f[i]);\n        }",
        "tests/test_oracle_pinning.py"),
    "fct_end": (
        "mpdata_oracle.c",
        "        } else {\n            fin_i2 = 0;\n            fout_i2 = 0;\n        }",
        "        } else {\n            fin_i2 = rmax2(L(0.), f[i]) - rmin2(L(0.), f[i]);\n"
        "            fout_i2 = rmax2(L(0.), f[i]) - rmin2(L(0.), f[i]);\n        }",
        "tests/test_oracle_pinning.py"),
    "wrf_x_both": (
        "std_oracle.c",
        "                } else if (u[FXI(i + 1, k, j)] < L(0.0)) {",
        "                } if (u[FXI(i + 1, k, j)] < L(0.0)) {",
        "tests/test_oracle_std.py"),
    "wrf_y_both": (
        "std_oracle.c",
        "                } else if (v[FYI(i, k, j + 1)] < L(0.0)) {",
        "                } if (v[FYI(i, k, j + 1)] < L(0.0)) {",
        "tests/test_oracle_std.py"),
    "wrf_w_above": (
        "std_oracle.c",
        "if (k > 0 && w[c - nx] > L(0.0)) {",
        "if (k > 0 && w[c] > L(0.0)) {",
        "tests/test_oracle_std.py"),
    "wrf_w_below": (
        "std_oracle.c",
        "if (k < nz - 1 && w[c] < L(0.0)) {",
        "if (k < nz - 1 && w[c - (k > 0 ? nx : 0)] < L(0.0)) {",
        "tests/test_oracle_std.py"),
}


@pytest.mark.skipif(shutil.which(os.environ.get("CC", "gcc")) is None, reason="no C compiler")
@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_fails_a_pinning_test(tmp_path, name):
    src, old, new, tests = MUTANTS[name]
    for d in ("oracle", "tests", "hicar_amd"):
        shutil.copytree(os.path.join(ROOT, d), tmp_path / d,
                        ignore=shutil.ignore_patterns("*.so", "_build", "__pycache__", "csrc"))
    shutil.copy(os.path.join(ROOT, "pytest.ini"), tmp_path / "pytest.ini")
    path = tmp_path / "oracle" / src
    text = path.read_text()
    assert text.count(old) == 1, f"mutation site of {name} not found exactly once"
    path.write_text(text.replace(old, new))
    subprocess.run(["make", "-s", "-C", str(tmp_path / "oracle")], check=True)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        tests], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=600)
    assert r.returncode == 1, f"mutant {name} survived the pinning tests:\n{r.stdout[-2000:]}"
    assert "FAILED" in r.stdout
