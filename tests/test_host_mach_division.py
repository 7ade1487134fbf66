"""The FP32 Mach number of the field statistics (csrc/diagnostics.hip,
mach_of) divides by c_s = float32(1/sqrt(3)) as fma(x, zh, x * zl) with
zh = RN(1/c_s), zl = RN(1/c_s - zh) (Brisebarre, Muller and Raina 2004).
The reference computes |u| / c_s with a float32 division (numpy), so the two
must agree bit for bit wherever the value is used.  Checked here on the host
with a C program (correctly rounded fmaf and division): every float of the
binades [1, 2) and [2, 4) -- the identity is invariant under scaling by
powers of two while x * zl stays normal -- and the top binades up to
FLT_MAX, plus the constants themselves against the kernel source."""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "pour-over-coffee-lbm_amd", "csrc", "diagnostics.hip")

PROGRAM = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static int same(float a, float b) { return memcmp(&a, &b, 4) == 0; }
int main(void) {
  volatile float y = (float)0.57735026918962573;
  long double inv = 1.0L / (long double)y;
  const float zh = (float)inv, zl = (float)(inv - (long double)zh);
  if (!same(zh, ZH) || !same(zl, ZL)) { printf("constants differ\n"); return 2; }
  long bad = 0;
  for (uint32_t b = 0x3f800000u; b < 0x40800000u; ++b) {  /* [1, 4) */
    float x; memcpy(&x, &b, 4);
    if (!same(x / y, fmaf(x, zh, x * zl))) ++bad;
  }
  for (uint32_t b = 0x7e000000u; b < 0x7f800000u; ++b) {  /* [2^125, FLT_MAX] */
    float x; memcpy(&x, &b, 4);
    if (!same(x / y, fmaf(x, zh, x * zl))) ++bad;
  }
  /* the smallest x of the claim: x * zl normal */
  float x0 = 0x1p-101f;
  if (!isnormal(x0 * zl)) ++bad;
  printf("%ld\n", bad);
  return bad != 0;
}
"""


def _kernel_constants():
    src = open(SRC).read()
    zh = re.search(r"constexpr float zh = ([0-9a-fx.p+-]+)f;", src).group(1)
    zl = re.search(r"constexpr float zl = ([0-9a-fx.p+-]+)f;", src).group(1)
    return zh, zl


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_constant_divisor_division_is_correctly_rounded(tmp_path):
    zh, zl = _kernel_constants()
    c = tmp_path / "mach.c"
    c.write_text(PROGRAM.replace("ZH", zh + "f").replace("ZL", zl + "f"))
    exe = tmp_path / "mach"
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(c), "-lm"])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "0"
