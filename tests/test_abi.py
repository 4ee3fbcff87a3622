"""The C-ABI libraries load and export every function include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import pathlib
import re

import pytest

from qtqp_amd import _native

INCLUDE = pathlib.Path(__file__).resolve().parents[1] / "include"
HEADER_TO_LIB = {"qtqp_kkt.h": "libqtqp_kkt.so", "qtqp_batch.h": "libqtqp_batch.so",
                 "qtqp_mlp.h": "libqtqp_mlp.so"}


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


def test_device_count_without_gpu_is_safe():
  lib = _native.kkt_lib()
  assert lib.qtqp_device_count() >= 0
  assert lib.qtqp_last_error() is not None
