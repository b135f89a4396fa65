"""The C-ABI library loads, exports every function include/sfci_gpu.h declares, and
rejects bad arguments before touching a GPU."""
import ctypes as C
import os

import pytest

from sfci_amd import gpu


def test_library_exports_every_declared_symbol():
    L = gpu.lib()
    names = gpu.declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_create_missing_file_is_io_error(tmp_path):
    h = C.c_void_p()
    rc = gpu.lib().sfci_create(str(tmp_path / "nope_IN").encode(), 24, 8, 0, C.byref(h))
    assert rc == -6 and not h.value


def test_create_bad_sizes_is_arg_error(make_system):
    s = make_system(4, 4, 0)
    h = C.c_void_p()
    assert gpu.lib().sfci_create(s["in_path"].encode(), 7, 4, 0, C.byref(h)) == -1     # odd n_so
    assert gpu.lib().sfci_create(s["in_path"].encode(), 66, 4, 0, C.byref(h)) == -1    # > 64 spin orbitals
    assert gpu.lib().sfci_create(s["in_path"].encode(), 8, 3, 0, C.byref(h)) == -1     # odd electrons


def test_null_arguments_rejected():
    L = gpu.lib()
    assert L.sfci_hv(None, None, None, 1, 0.0, 1.0, None) == -1
    assert L.sfci_set_cut(None, 1e-5) == -1
    assert L.sfci_vec_len(None) == -1
