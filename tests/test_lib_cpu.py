"""CPU-side checks of the C ABI: the library loads and exports every entry point
that include/gsr.h declares; the Python binding names them all.  No compute."""
import ctypes
import os
import re

from gsr import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "gsr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gsr_[a-z_]+)\s*\(", text)))


def test_header_declares_the_operator_surface():
    names = declared()
    for op in ("gsr_project", "gsr_bin_sort", "gsr_forward", "gsr_loss_grad", "gsr_backward", "gsr_shape",
               "gsr_adam", "gsr_stats", "gsr_batch_step"):
        assert op in names


def test_library_exports_every_declared_symbol():
    so = ctypes.CDLL(lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(so, n)]
    assert not missing, missing
    assert sorted(lib.EXPORTS) == declared()


def test_binding_signatures_load():
    l = lib.load()
    for n in lib.EXPORTS:
        assert getattr(l, n).argtypes is not None or getattr(l, n).restype is None


def test_struct_layouts():
    assert ctypes.sizeof(lib.Camera) == 80
    assert ctypes.sizeof(lib.Step) == 32
    assert lib.Camera.count.offset == 72 and lib.Camera.pixels.offset == 76
