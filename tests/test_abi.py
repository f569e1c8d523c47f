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


def test_block_structure_arguments_are_checked():
    """The *_blocks entry points refuse segment starts that do not begin at 0
    and increase below n before touching the device."""
    import ctypes as C
    from openmeeg_amd import _lib
    L = _lib.lib()
    hfill = np.zeros((3, 3), dtype=np.uint8)
    info = C.c_int32()
    dummy = C.c_void_p(16)               # never dereferenced: the check fails first
    hp = hfill.ctypes.data_as(C.c_void_p)
    for starts in ([0, 20, 10], [1, 10, 20], [0, 10, 30]):
        st = np.array(starts, dtype=np.int64)
        sp = st.ctypes.data_as(C.c_void_p)
        assert L.omh_sym_factor_f32_split_blocks(dummy, 30, 30, dummy, 3, sp, hp, C.byref(info), C.c_void_p(0)) == 2
        assert b"segment" in L.omh_last_error()
        assert L.omh_sym_solve_f64_blocks(dummy, 30, 30, dummy, dummy, 30, 4, 3, sp, hp, C.c_void_p(0)) == 2
        it, rel = C.c_int32(), C.c_double()
        assert L.omh_mixed_solve_blocks(dummy, 30, 30, dummy, 30, dummy, dummy, 30, dummy, 30, 4, 5, 1e-8, 3, sp, hp,
                                        C.byref(it), C.byref(rel), C.c_void_p(0)) == 2
        assert b"segment" in L.omh_last_error()


def test_head1_block_pattern_matches_golden_headmat():
    """Geometry.factor_segments' block pattern of the HeadMat against the
    reference's HeadMat of Head1 (tests/golden/head1/hm.npy): every block it
    declares zero is exactly zero there, every block it declares non-zero
    holds entries."""
    import openmeeg_amd as om
    g = om.read_geometry(os.path.join(GOLDEN, "head1", "Head1.geom"), os.path.join(GOLDEN, "head1", "Head1.cond"))
    from oracle import oracle as O
    n = g.size
    hm = O.unpack(np.load(os.path.join(GOLDEN, "head1", "hm.npy")), n)   # packed upper -> dense
    p = g.tri_first_index(np.arange(n))
    Htf = np.empty_like(hm)
    Htf[np.ix_(p, p)] = hm
    starts, hfill = g.factor_segments()
    ends = list(starts[1:]) + [n]
    ns = len(starts)
    assert hfill.shape == (ns, ns) and ns >= 3
    for s in range(ns):
        for t in range(s):
            blk = Htf[starts[s]:ends[s], starts[t]:ends[t]]
            assert bool(np.any(blk != 0)) == bool(hfill[s, t]), (s, t)


def test_source_mesh_handle_on_the_host():
    """Mesh(.tri) and Mesh(vertices, triangles) build the handle on the host
    (no device needed); out-of-range triangle indices and empty meshes are
    refused with invalid-argument status."""
    import ctypes as C
    import openmeeg_amd as om
    from openmeeg_amd import _lib
    m = om.Mesh(os.path.join(GOLDEN, "head1", "Head1.tri"))
    assert m.vertices().size() == 42 and m.nb_triangles() == 80
    sz = (C.c_int64 * 2)()
    _lib.check(_lib.lib().omh_mesh_sizes(m._h, sz))
    assert list(sz) == [42, 80]
    P = np.eye(3)
    with pytest.raises(ValueError, match="out of range"):
        om.Mesh(P, [[0, 1, 3]])
    with pytest.raises(ValueError, match="empty"):
        om.Mesh(np.zeros((0, 3)), np.zeros((0, 3), dtype=np.int64))


def test_refinement_stopping_rule():
    """omh_refine_state (the stop of omh_mixed_solve*): 1 converged, 2
    stalled, 0 continue.  Divergence and stagnation above the FP64 floor are
    failures, not convergence; stagnation at the floor is convergence."""
    from openmeeg_amd import _lib
    st = _lib.lib().omh_refine_state
    tol = 1e-8

    def run(seq, tol=tol):
        prev = 2.0
        for it, rel in enumerate(seq):
            s = st(it, rel, prev, tol)
            if s:
                return s, it
            prev = rel
        return 0, len(seq)

    assert run([3e-5, 6e-8, 2.7e-10]) == (1, 1)              # contracting: 6e-8 x 2e-3 <= 1e-8
    assert run([3e-5, 6e-6, 1e-6, 2e-7]) == (0, 4)           # estimates 1.2e-6, 1.7e-7, 4e-8: continue
    assert run([3e-5, 6e-6, 1e-6, 2e-7, 1e-9]) == (1, 4)
    assert run([1e-9]) == (1, 0)                             # the update itself below tol
    assert run([1e-2, 5e-2, 1e-1, 1.0]) == (2, 2)            # rising: divergence is a failure
    assert run([1e-3, 9e-4, 8.5e-4]) == (2, 2)               # stalled far above tol
    assert run([1e-3, 9e-4, 8.5e-4], tol=1e-16) == (2, 2)
    assert run([1e-10, 2e-12, 1.5e-12], tol=1e-14) == (1, 2)  # stalled at the FP64 floor
    assert run([1e-10, 2e-12, 1.5e-12], tol=1e-30) == (2, 2)  # ... but not within 1e3 tol of it
    assert run([3e-2, 1e-2, 3.3e-3, 1.1e-3]) == (0, 4)        # slow contraction continues (until max_iter)
    assert st(2, float("nan"), 1e-3, tol) == 2
    assert st(1, 5e-9, 1e-3, tol) == 1
    assert st(1, 2e-3, 1e-3, tol) == 0                        # a rise on the second correction is allowed once
