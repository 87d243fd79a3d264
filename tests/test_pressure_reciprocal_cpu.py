"""The contact-list collide mode takes the pressure quotient len / (4 pi r^2)
from the particle's rounded reciprocal y = RN(1/den) and one fused
correction, q' = RN(q + RN(len - q den) y) with q = RN(len y) (Markstein's
theorem).  This checks that sequence against correctly rounded division on
random float lengths (zero, subnormal and normal up to FLT_MAX) and radii
from 2^-40 to 2^40, bit for bit, with the host C library's fma; a variant
without the correction is shown to fail, so the check has teeth."""
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parent / "pressure_reciprocal_check.c"


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


def test_reciprocal_quotient_is_correctly_rounded(tmp_path):
    exe = build(tmp_path, SRC.read_text())
    assert mismatches(exe, 20_000_000) == 0


def test_uncorrected_quotient_fails(tmp_path):
    text = SRC.read_text().replace("return isfinite(r) ? fma(r, y, q) : a / den;", "(void)r; return q;")
    exe = build(tmp_path, text)
    assert mismatches(exe, 200_000) > 1000
