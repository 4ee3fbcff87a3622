"""The C-ABI libraries load and export every function include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import pathlib
import re
import shutil
import subprocess

import pytest

from qtqp_amd import _native, batch

INCLUDE = pathlib.Path(__file__).resolve().parents[1] / "include"
HEADER_TO_LIB = {"qtqp_kkt.h": "libqtqp_kkt.so", "qtqp_ipm.h": "libqtqp_kkt.so",
                 "qtqp_batch.h": "libqtqp_batch.so", "qtqp_mlp.h": "libqtqp_mlp.so",
                 "qtqp_sparse.h": "libqtqp_sparse.so"}


def declared(header: pathlib.Path):
  text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
  return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w]+\s*\**\s*(qtqp_\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("header", sorted(p.name for p in INCLUDE.glob("*.h")))
def test_library_exports_every_declared_symbol(header):
  names = declared(INCLUDE / header)
  assert names, header
  path = _native.lib_path(HEADER_TO_LIB[header])
  assert path.exists(), f"{path} not built"
  lib = ctypes.CDLL(str(path))
  missing = [n for n in names if not hasattr(lib, n)]
  assert not missing, missing


def test_kkt_bindings_cover_header():
  assert set(declared(INCLUDE / "qtqp_kkt.h")) == set(_native.KKT_SYMBOLS)
  assert set(declared(INCLUDE / "qtqp_ipm.h")) == set(_native.IPM_SYMBOLS)
  assert set(declared(INCLUDE / "qtqp_sparse.h")) == set(_native.SPARSE_SYMBOLS)


def _c_layout(tmp_path, header, struct, fields):
  """sizeof and field offsets of a C struct, from gcc on the header itself."""
  prog = tmp_path / "layout.c"
  body = "".join(f'  printf("%zu\\n", offsetof({struct}, {f}));\n' for f in fields)
  prog.write_text(f'#include <stddef.h>\n#include <stdio.h>\n#include "{header}"\n'
                  f'int main(void) {{\n  printf("%zu\\n", sizeof({struct}));\n{body}  return 0;\n}}\n')
  exe = tmp_path / "layout"
  subprocess.run(["gcc", f"-I{INCLUDE}", str(prog), "-o", str(exe)], check=True)
  out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
  return [int(v) for v in out]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("cls,struct,header", [
    (_native.IpmOpts, "qtqp_ipm_opts", "qtqp_ipm.h"),
    (_native.IpmRecord, "qtqp_ipm_record", "qtqp_ipm.h"),
    (_native.IrStats, "qtqp_ir_stats", "qtqp_kkt.h"),
    (_native.KktInfo, "qtqp_kkt_info", "qtqp_kkt.h"),
    (_native.SparseInfo, "qtqp_sparse_info", "qtqp_sparse.h"),
    (batch.BatchOpts, "qtqp_batch_opts", "qtqp_batch.h"),
    (batch.BatchCounters, "qtqp_batch_counters", "qtqp_batch.h")])
def test_ctypes_structs_match_c_layout(tmp_path, cls, struct, header):
  names = [f for f, _ in cls._fields_]
  c = _c_layout(tmp_path, header, struct, names)
  assert c[0] == ctypes.sizeof(cls)
  assert c[1:] == [getattr(cls, f).offset for f in names]


def test_device_count_without_gpu_is_safe():
  lib = _native.kkt_lib()
  assert lib.qtqp_device_count() >= 0
  assert lib.qtqp_last_error() is not None
