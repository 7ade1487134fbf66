"""Build oracle/liboracle_lbm.so from lbm_oracle.c (test infrastructure).

The one source is compiled twice, REAL=double (symbols *_f64) and
REAL=float (symbols *_f32), with OpenMP and without floating-point
contraction, then linked into one shared library beside this file.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "lbm_oracle.c"
LIB = HERE / "liboracle_lbm.so"
BUILD = HERE / "_build"

CFLAGS = ["-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c11", "-Wall", "-Wno-unused-function"]


def build() -> Path:
    cc = os.environ.get("CC") or shutil.which("gcc") or "cc"
    BUILD.mkdir(exist_ok=True)
    objs = []
    for real, sfx in (("double", "f64"), ("float", "f32")):
        obj = BUILD / ("lbm_oracle_%s.o" % sfx)
        cmd = [cc, *CFLAGS, "-DREAL=%s" % real, "-DSFX(n)=n##_%s" % sfx, "-c", str(SRC), "-o", str(obj)]
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    subprocess.run([cc, "-shared", "-fopenmp", "-o", str(tmp), *objs, "-lm"], check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
