"""Batched EEG gains: many heads of uneven size in one pass.

``BatchedGain([(geometry, dipoles, electrodes), ...], precision)`` wraps the
library's batch plan (omh_batch_*): every head's HeadMat and source matrix
are assembled (the heads spread over four streams), and the dense algebra of
all heads (factorisation, adjoint solve with refinement, gain product) runs
as grouped launches in lock step.  After a first eager pass,
``run(graph=True)`` replays the whole pass as one captured HIP graph;
``tune()`` picks the trailing update by timing the candidates on the
batch's own shapes: FP32 products on 64- or 128-wide tiles and grid caps,
or ("mixed") the split-FP16 update, the default before tuning.  The reference computes
one head per call (gain.h:78-89); there is no batched counterpart.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib
from .api import _dipoles_array, _electrodes_array, _stream, _t


class _DeviceArray:
    """__cuda_array_interface__ view of a library-owned device buffer."""

    def __init__(self, ptr, shape):
        self.__cuda_array_interface__ = dict(shape=shape, typestr="<f8", data=(ptr, False), version=3, strides=None)


class BatchedGain:
    PRECISIONS = {"fp64": 0, "mixed": 1}

    def __init__(self, heads, precision="mixed"):
        if precision not in self.PRECISIONS:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        _t()
        self.precision = precision
        self._geoms = [h[0] for h in heads]          # keep the geometries alive
        dips = [np.ascontiguousarray(_dipoles_array(h[1])) for h in heads]
        els = [np.ascontiguousarray(_electrodes_array(h[2])) for h in heads]
        nh = len(heads)
        if nh == 0:
            raise ValueError("a batch needs at least one head")
        G = (C.c_void_p * nh)(*[g._h.value for g in self._geoms])
        D = (C.c_void_p * nh)(*[d.ctypes.data for d in dips])
        E = (C.c_void_p * nh)(*[e.ctypes.data for e in els])
        nd = (C.c_int64 * nh)(*[len(d) for d in dips])
        ne = (C.c_int64 * nh)(*[len(e) for e in els])
        self._h = C.c_void_p()
        check(lib().omh_batch_create(nh, G, D, nd, E, ne, self.PRECISIONS[precision], C.byref(self._h)))
        self.n_heads = nh
        self.used_graph = False

    def __del__(self):
        h = getattr(self, "_h", None)
        L = getattr(_lib, "_lib", None) if _lib is not None else None
        if h and h.value and L is not None:
            L.omh_batch_free(h)
            self._h = None

    def run(self, graph=True):
        """One pass; returns True when it replayed the captured graph."""
        used = C.c_int32()
        check(lib().omh_batch_run(self._h, 1 if graph else 0, C.byref(used), _stream()))
        self.used_graph = bool(used.value)
        return self.used_graph

    def tune(self):
        """Time the trailing-update candidates; returns [(tile, grid_cap,
        split16, seconds)] and keeps the fastest."""
        buf = (C.c_double * 128)()
        n = C.c_int32()
        check(lib().omh_batch_tune(self._h, buf, 32, C.byref(n), _stream()))
        return [(int(buf[4 * i]), int(buf[4 * i + 1]), bool(buf[4 * i + 2]), buf[4 * i + 3])
                for i in range(min(n.value, 32))]

    def config(self):
        rs, tile, split = C.c_int32(), C.c_int32(), C.c_int32()
        cap, launches, nodes = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().omh_batch_config(self._h, C.byref(rs), C.byref(tile), C.byref(cap), C.byref(launches),
                                     C.byref(nodes), C.byref(split)))
        return dict(refine_steps=rs.value, tile=tile.value, grid_cap=cap.value, split16=bool(split.value),
                    grouped_launches=launches.value, graph_nodes=nodes.value)

    def gain_device(self, h):
        """Head h's gain as a torch tensor (n_el x n_dip) viewing the batch's
        buffer (valid while the batch lives; the next pass overwrites it)."""
        torch = _t()
        p = C.c_void_p()
        ne, nd = C.c_int64(), C.c_int64()
        check(lib().omh_batch_gain(self._h, h, C.byref(p), C.byref(ne), C.byref(nd)))
        t = torch.as_tensor(_DeviceArray(p.value, (nd.value, ne.value)), device="cuda")
        return t.t()

    def gain(self, h):
        """Head h's gain as a numpy array (n_el x n_dip)."""
        _t().cuda.current_stream().synchronize()
        return self.gain_device(h).cpu().numpy()

    def stats(self, h):
        info, rel = C.c_int32(), C.c_double()
        check(lib().omh_batch_stats(self._h, h, C.byref(info), C.byref(rel), _stream()))
        return dict(factor_info=info.value, rel_update=rel.value)
