"""Builds the HIP shared libraries of qtqp_amd in-tree for gfx950.

Each library is compiled by hipcc from its sources under qtqp_amd/csrc into
qtqp_amd/_lib/.  Usage: ``python -m qtqp_amd.build`` (or
``__graft_entry__.build()``).
"""

from __future__ import annotations

import concurrent.futures
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
INCLUDE = ROOT / "include"

# name -> source files
LIBRARIES = {
    "libqtqp_kkt.so": ("kkt.hip", "ipm.hip"),
    "libqtqp_batch.so": ("batch.hip",),
    "libqtqp_mlp.so": ("mlp.hip",),
    "libqtqp_sparse.so": ("sparse.hip",),
}

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # Packed FP32 VALU forms are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Keep the exec-mask restore at the end of every divergent region.  With
    # the backend's default, the restore of a region nested at the end of
    # another is dropped in favour of the outer one; a register copy that
    # the allocator later places between the two then runs with an empty exec
    # mask, and the value it was meant to restore keeps whatever the region
    # wrote into its register (seen here: a refinement loop's residual norm
    # replaced by the last element a copy loop moved).  tests/test_codegen.py
    # scans the built code for that pattern.
    "-mllvm", "-amdgpu-remove-redundant-endcf=0",
    "-Wno-unused-result",
]


def hipcc() -> str:
  for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
    if cand and pathlib.Path(cand).exists():
      return cand
  raise RuntimeError("hipcc not found; the HIP libraries cannot be built")


def _build_one(lib: str, srcs, verbose: bool):
  paths = [CSRC / src for src in srcs]
  if not all(p.exists() for p in paths):
    return None
  out = LIBDIR / lib
  tmp = out.with_suffix(".so.tmp")
  cmd = [hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", *map(str, paths), "-o", str(tmp)]
  if verbose:
    print("[build]", " ".join(cmd), flush=True)
  res = subprocess.run(cmd, capture_output=True, text=True)
  if res.returncode != 0:
    sys.stderr.write(res.stdout + res.stderr)
    raise RuntimeError(f"hipcc failed for {', '.join(srcs)}")
  os.replace(tmp, out)
  return out


def build(names=None, verbose: bool = True) -> list[pathlib.Path]:
  """Compiles the libraries (all, or those in `names`), one hipcc process each,
  in parallel."""
  LIBDIR.mkdir(exist_ok=True)
  todo = [(lib, srcs) for lib, srcs in LIBRARIES.items() if not names or lib in names]
  with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, len(todo))) as ex:
    outs = list(ex.map(lambda t: _build_one(t[0], t[1], verbose), todo))
  return [o for o in outs if o is not None]


if __name__ == "__main__":
  build(names=sys.argv[1:] or None)
