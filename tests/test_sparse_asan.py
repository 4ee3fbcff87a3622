"""The sparse symbolic analysis under AddressSanitizer (host code only, no GPU):
tests/asan_sparse_analyze.cpp is built with the sanitizer on its host side and
run on random KKT patterns with each ordering; any out-of-bounds access, use
after free or leak fails the test."""
import os
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_symbolic_analysis_under_address_sanitizer(tmp_path):
  exe = tmp_path / "asan_sparse_analyze"
  build = [HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer",
           f"-I{ROOT / 'include'}", f"-I{ROOT / 'qtqp_amd' / 'csrc'}",
           str(ROOT / "tests" / "asan_sparse_analyze.cpp"), "-o", str(exe)]
  subprocess.run(build, check=True, capture_output=True, timeout=600)
  env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0")
  run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
  assert "AddressSanitizer" not in run.stderr, run.stderr[-4000:]
  assert "LeakSanitizer" not in run.stderr, run.stderr[-4000:]
  assert run.returncode == 0, run.stdout[-2000:] + run.stderr[-2000:]
  assert run.stdout.count("ok=1") == 18
