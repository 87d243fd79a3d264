"""The contact-list collide mode takes the pressure quotient len / (4 pi r^2)
from the particle's rounded reciprocal y = RN(1/den) and two fused
corrections, q <- RN(q + RN(len - q den) y) from q = RN(len y): the first
makes q faithful, the second correctly rounded (Markstein's theorem, which
needs a faithful q).  This checks that sequence against correctly rounded
division on random float lengths (zero, subnormal and normal up to FLT_MAX)
and radii from 2^-40 to 2^40, bit for bit, with the host C library's fma; a
variant without the corrections is shown to fail, so the check has teeth, and
the uncorrected quotient is shown not to be faithful on some inputs, which is
why one correction is not enough for the theorem."""
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


def mismatches(exe, count, *extra):
    out = subprocess.run([str(exe), str(count), *extra], check=True, capture_output=True, text=True).stdout
    return int(out.strip().splitlines()[-1])


def test_reciprocal_quotient_is_correctly_rounded(tmp_path):
    exe = build(tmp_path, SRC.read_text())
    assert mismatches(exe, 20_000_000) == 0


def test_uncorrected_quotient_fails(tmp_path):
    text = SRC.read_text()
    corrected = "  q = fma(r, y, q);\n  return fma(fma(-q, den, a), y, q);"
    assert corrected in text
    exe = build(tmp_path, text.replace(corrected, "  return q;"))
    assert mismatches(exe, 200_000) > 1000


def test_uncorrected_quotient_is_not_always_faithful(tmp_path):
    exe = build(tmp_path, SRC.read_text())
    assert mismatches(exe, 1_000_000, "faithful") > 100
