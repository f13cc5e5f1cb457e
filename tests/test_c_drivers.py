"""Drop-in at the C level, on the CPU: programs written against the
reference's API compile against include/hydrofv/hydro.h and link against
libhydrofv_amd.so with gcc, unchanged -- the reference's own example drivers
(examples/*/*.c, where the reference sources are present) and this
repository's C driver (tests/native/kh_driver.c, which the GPU tests run)."""
from __future__ import annotations

import glob
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXAMPLES = "/root/reference/examples"
LIB_DIR = os.path.join(ROOT, "hydrofv_amd")


def compile_and_link(src, out):
    cmd = ["gcc", "-std=gnu99", "-O1", "-Werror=implicit-function-declaration", "-I",
           os.path.join(ROOT, "include", "hydrofv"), src, "-L", LIB_DIR, "-lhydrofv_amd",
           "-Wl,-rpath," + LIB_DIR, "-Wl,--no-undefined", "-lm", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stderr}"


def test_own_driver_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(LIB_DIR, "libhydrofv_amd.so")):
        pytest.fail("libhydrofv_amd.so is not built")
    compile_and_link(os.path.join(ROOT, "tests", "native", "kh_driver.c"), str(tmp_path / "kh_driver"))


@pytest.mark.skipif(not os.path.isdir(EXAMPLES), reason="reference examples not present")
def test_reference_examples_compile_and_link(tmp_path):
    """Every example driver of the reference, compiled where it lies (never
    run here; the one prebuilt binary among them is not touched)."""
    sources = sorted(glob.glob(os.path.join(EXAMPLES, "*", "*.c")))
    assert len(sources) >= 8
    for src in sources:
        compile_and_link(src, str(tmp_path / os.path.basename(src)[:-2]))
