"""The step kernels decide the shear law's Coulomb test, sqrtf(x) < m
(src/particles_kernel.cuh:295), from x against RN(m m) (1 -/+ 2^-20)
where that is certain and take the correctly rounded root only in between.
This checks that decision against sqrtf(x) < m with the host C library, on
random m and on every float within 64 ulps of m*m, of the square of the
midpoint below m and of both thresholds; a variant with the margins
dropped is shown to fail, so the check has teeth."""
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parent / "sqrt_below_check.c"


def build(tmp_path, src_text):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler on this host")
    src = tmp_path / "check.c"
    src.write_text(src_text)
    exe = tmp_path / "check"
    subprocess.run([cc, "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    return exe


def mismatches(exe, count):
    out = subprocess.run([str(exe), str(count)], check=True, capture_output=True, text=True).stdout
    return int(out.strip().splitlines()[-1])


def test_fast_coulomb_test_matches_the_root(tmp_path):
    assert mismatches(build(tmp_path, SRC.read_text()), 200_000) == 0


def test_fast_coulomb_test_without_margin_fails(tmp_path):
    text = SRC.read_text()
    for sign in "-+":
        old = f"t * (1.0f {sign} 0x1p-20f)"
        assert old in text
        text = text.replace(old, "t")
    assert mismatches(build(tmp_path, text), 20_000) > 100
