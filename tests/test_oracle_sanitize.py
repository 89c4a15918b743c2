"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/sanitize_main.c runs every oracle function (octree codes, the whole solve, iso
value, dual contouring in count and fill mode, and the grid kernels on odd and even sizes)
on small inputs; built with -fsanitize=address,undefined and no recovery, any out-of-bounds
access, leak or undefined operation ends the run with a non-zero status.
"""

from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_is_clean_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    exe = tmp_path / "oracle_sanitize"
    subprocess.run(["gcc", "-O1", "-g", "-std=c11", "-ffp-contract=off",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", os.path.join(ORACLE, "sanitize_main.c"),
                    os.path.join(ORACLE, "mpr_oracle.c"), "-I", ORACLE, "-lm", "-o", str(exe)],
                   check=True, capture_output=True, timeout=120)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, env=env)
    assert run.returncode == 0, run.stdout + run.stderr
    assert run.stdout.startswith("sanitize ok")
