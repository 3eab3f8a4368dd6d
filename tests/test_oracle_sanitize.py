"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/sanitize_driver.c calls every oracle entry point on odd and degenerate
shapes (axes of 1 and 2 nodes, many substrates, masked and repeated Dirichlet
nodes, bulk fields, stacked agents, an agent outside the domain, the 2-D solver,
gradients); built with -fsanitize=address,undefined -fno-sanitize-recover=all,
any out-of-bounds access, use after free or undefined operation aborts it.
Host code only: no GPU is involved.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not installed")
def test_oracle_runs_clean_under_address_and_undefined_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_sanitize")
    cmd = ["gcc", "-O1", "-g", "-std=c11", "-ffp-contract=off", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           os.path.join(ORACLE, "bfvm_oracle.c"), os.path.join(ORACLE, "sanitize_driver.c"), "-lm", "-o", exe]
    build = subprocess.run(cmd, capture_output=True, text=True)
    if build.returncode != 0 and "sanitizer" in (build.stderr + build.stdout).lower():
        pytest.skip("the sanitizer runtimes are not installed: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "runtime error" not in run.stderr, run.stderr
    assert "18 cases, ok" in run.stdout, run.stdout
