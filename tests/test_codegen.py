"""The device code the build produces has no vector instruction that runs
with an empty exec mask after a divergent loop.

A divergent loop on AMDGPU ends when every lane has left it (s_andn2 exec;
s_cbranch_execnz back to its head); the exec mask is then empty until the
enclosing region restores it.  A vector instruction placed there (a register
copy the allocator inserted to restore a value live across the loop) does
nothing, and the value keeps what the loop wrote into its register.  The
backend's removal of "redundant" region ends produced exactly that in the
batched engine's refinement loop; build.py turns it off.  This test compiles
every library's device code to assembly with the build's flags and scans it
for vector instructions between such a loop exit and the next write of exec.
Lane-indexed moves (v_readlane / v_writelane) ignore exec and are allowed.
"""
import concurrent.futures
import pathlib
import re
import subprocess

import pytest

from qtqp_amd import build

SOURCES = [src for srcs in build.LIBRARIES.values() for src in srcs]


def _asm(src: str, out: pathlib.Path) -> str:
  flags = [f for f in build.HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
  cmd = [build.hipcc(), *flags, f"-I{build.INCLUDE}", f"-I{build.CSRC}", "--cuda-device-only",
         "-S", str(build.CSRC / src), "-o", str(out)]
  res = subprocess.run(cmd, capture_output=True, text=True)
  assert res.returncode == 0, res.stderr[-2000:]
  return out.read_text()


EXEC_WRITE = re.compile(r"s_\w+\s+exec\b|s_\w+_saveexec_b64|s_(c?branch|setpc_b64|endpgm)")
VECTOR = re.compile(r"(v_|flat_|global_|buffer_|ds_|scratch_)")
LANE_OP = re.compile(r"v_(readlane|writelane|readfirstlane)_b32")


def empty_exec_vector_ops(text: str):
  """(line, instruction) of vector ops right after a divergent loop's exit."""
  lines = text.splitlines()
  hits = []
  for i, line in enumerate(lines):
    if not re.match(r"\s*s_cbranch_execnz\s+\.LBB", line):
      continue
    prev = [l.strip() for l in lines[max(0, i - 3):i]]
    if not any(re.match(r"s_andn2_b64\s+exec, exec,", p) for p in prev):
      continue  # not the back edge of a divergent loop
    for j in range(i + 1, len(lines)):
      t = lines[j].strip()
      if not t or t.startswith(";") or re.match(r"\.LBB\w+:", t):
        continue
      if EXEC_WRITE.match(t):
        break
      if VECTOR.match(t) and not LANE_OP.match(t):
        hits.append((j + 1, t))
        break
  return hits


@pytest.mark.skipif(not pathlib.Path("/opt/rocm/bin/hipcc").exists(), reason="needs hipcc")
def test_no_vector_ops_with_empty_exec_after_divergent_loops(tmp_path):
  with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
    texts = list(ex.map(lambda s: _asm(s, tmp_path / (s + ".s")), SOURCES))
  found = {src: empty_exec_vector_ops(t) for src, t in zip(SOURCES, texts)}
  assert all(not h for h in found.values()), found
  assert all("s_andn2_b64 exec, exec," in t for t in texts)  # the scan saw divergent loops


def test_scanner_finds_the_pattern():
  bad = """
.LBB22_72:
	flat_load_dwordx2 v[8:9], v[2:3]
	s_andn2_b64 exec, exec, s[4:5]
	s_cbranch_execnz .LBB22_72
.LBB22_73:
	v_mov_b64_e32 v[8:9], v[10:11]
.LBB22_74:
	s_or_b64 exec, exec, s[2:3]
"""
  good = bad.replace(".LBB22_73:\n", ".LBB22_73:\n\ts_or_b64 exec, exec, s[4:5]\n")
  assert empty_exec_vector_ops(bad) == [(7, "v_mov_b64_e32 v[8:9], v[10:11]")]
  assert empty_exec_vector_ops(good) == []
