"""The C-ABI library loads and exports every symbol include/*.h declares (no GPU needed)."""
import ctypes as C
import glob
import os
import re

import pytest

import alpine_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(apg_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_headers_declare_symbols():
    names = declared_symbols()
    assert "apg_intersect" in names and "apg_occluded" in names and len(names) >= 20


@pytest.mark.parametrize("name", declared_symbols())
def test_library_exports(name):
    L = C.CDLL(alpine_amd.LIB_PATH)
    assert hasattr(L, name), f"{name} declared in include/ but not exported"


def test_python_binding_lists_every_symbol():
    assert sorted(alpine_amd.ABI_SYMBOLS + alpine_amd.DEBUG_SYMBOLS) == declared_symbols()


def test_create_without_gpu_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(alpine_amd.ApgError) as e:
        alpine_amd.GpuAccelerator()
    assert e.value.code == -2
    assert b"not available" in alpine_amd.lib().apg_last_error(None)


def test_null_context_is_rejected():
    L = alpine_amd.lib()
    assert L.apg_intersect(None, None, 0, None, None) == -1
    assert L.apg_scene_build(None, None) == -1
