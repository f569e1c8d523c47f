"""The C ABI library loads and exports every symbol include/openmeeg_amd.h
declares; host-only entry points (geometry, error paths) work without a GPU."""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


def _declared():
    txt = open(os.path.join(ROOT, "include", "openmeeg_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(omh_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from openmeeg_amd import _lib
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.symbols())


def test_error_status_and_message():
    import ctypes as C
    from openmeeg_amd import _lib
    h = C.c_void_p()
    rc = _lib.lib().omh_geometry_load(b"/nonexistent.geom", b"/nonexistent.cond", C.byref(h))
    assert rc != 0
    assert b"cannot open" in _lib.lib().omh_last_error()
    with pytest.raises(RuntimeError):
        _lib.check(rc)


def test_workspace_size_query():
    from openmeeg_amd import _lib
    assert _lib.lib().omh_factor_workspace_bytes(1000, 8) > 8 * 1000 * 256
    assert _lib.lib().omh_factor_workspace_bytes(1000, 3) == -1


def test_gpu_entry_points_refuse_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import openmeeg_amd as om
    g = om.read_geometry(os.path.join(GOLDEN, "head1", "Head1.geom"), os.path.join(GOLDEN, "head1", "Head1.cond"))
    with pytest.raises(RuntimeError, match="GPU"):
        om.HeadMat(g)


def test_block_structure_must_be_closed():
    """omh_sym_factor_f32_split_blocks refuses a segment pattern that is not
    closed under elimination before touching the device (two segments reached
    from the same earlier segment must be coupled below it)."""
    import ctypes as C
    from openmeeg_amd import _lib
    starts = np.array([0, 10, 20], dtype=np.int64)
    fill = np.zeros((3, 3), dtype=np.uint8)
    fill[1, 0] = fill[2, 0] = 1          # 1 and 2 both below 0 ...
    info = C.c_int32()
    dummy = C.c_void_p(16)               # never dereferenced: the check fails first
    rc = _lib.lib().omh_sym_factor_f32_split_blocks(dummy, 30, 30, dummy, 3, starts.ctypes.data_as(C.c_void_p),
                                                     fill.ctypes.data_as(C.c_void_p), C.byref(info), C.c_void_p(0))
    assert rc == 2                       # ... but 2 is not coupled to 1
    assert b"closed" in _lib.lib().omh_last_error()
