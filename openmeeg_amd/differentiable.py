"""Differentiable EEG forward model.

``EEGForward(geom, electrodes).gain(positions, moments, conductivity)``
returns the EEG gain matrix G = E H^-1 D (gain.h:78-89: E the Head2EEGMat
interpolation, H the HeadMat, D the dipole source matrix) as a torch tensor
that carries gradients with respect to the dipole positions, the dipole
moments and the domain conductivities.  The reference has no derivatives;
this is the adjoint of its forward path.

Adjoint relations, with Y = H^-1 E^T (the solve the forward already does)
and Gbar = dL/dG:
  * dL/dD = W = Y Gbar; the position / moment gradients contract W with the
    derivatives of the two source integrands (dipole.h:29-46) on the
    device (omh_dip_source_adjoint, one fused kernel).
  * dL/dH = -Y Z^T with Z = H^-1 (D Gbar^T) (n_electrodes right-hand sides).
    H depends on the conductivities through the pair coefficients
    (assembleHeadMat.cpp:92-97): the S blocks scale with sum 1/sigma, the N
    blocks with sum sigma over the pair's common domains, and the deflation
    constant (assembleHeadMat.cpp:28-60) with the N block that sets it.  Each
    block's share <Y_R, H_RC Z_C> is one GEMM on the stored H.
  * The triangle (potential) rows of D carry 1/sigma of the dipole's domain
    (assembleSourceMat.cpp:84-99).

The source matrix uses the non-adaptive rule of the given order, so the
gradients are exact derivatives of the computed gain.  Everything runs in
the HIP library; PyTorch holds the buffers.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib, ORDER_TRIANGLES_FIRST
from .api import (Head2EEGMat, Integrator, _dev_f64, _eeg_rhs, _head_mat_device, _ptr, _stream, _t, factor)


def _gemm(ta, tb, m, n, k, A, lda, B, ldb, Cout, ldc, a_off=0, b_off=0, alpha=1.0, beta=0.0):
    """omh_gemm_f64 on (offset) views of flat float64 device tensors."""
    pa = C.c_void_p(A.data_ptr() + 8 * a_off)
    pb = C.c_void_p(B.data_ptr() + 8 * b_off)
    check(lib().omh_gemm_f64(ta, tb, m, n, k, alpha, pa, lda, pb, ldb, beta, _ptr(Cout), ldc, _stream()))


class EEGForward:
    """EEG forward model on one geometry and electrode set, differentiable in
    the dipoles and the conductivities.  precision 'mixed' (FP32 factor with
    FP64 refinement), 'mixed16' (split-FP16 factor with FP64 refinement) or
    'fp64'."""

    def __init__(self, geom, electrodes, precision="mixed", order=3, max_refine=10, refine_tol=1e-12):
        self.geom = geom
        self.precision = precision
        self.order = order
        self.max_refine, self.refine_tol = max_refine, refine_tol
        self.h2em = Head2EEGMat(geom, electrodes)       # conductivity independent
        self.ne = self.h2em.nlin()
        self._E = _eeg_rhs(geom, self.h2em)
        self._sigma = geom.conductivities()
        self._layout()

    def _layout(self):
        g = self.geom
        n, nv = g.size, g.n_vertex_unknowns
        nt = n - nv
        self.meshes = [g.mesh_info(m) for m in range(g.n_meshes)]
        self.pairs = g.mesh_pairs()
        ranges = []
        for M in self.meshes:
            if M["isolated"]:
                continue
            if M["vert_first"] < 0:
                raise NotImplementedError("conductivity gradients need each mesh's vertex unknowns contiguous")
            ranges.append((M["vert_first"], M["vert_end"]))
        ranges.sort()
        for (a0, a1), (b0, b1) in zip(ranges[:-1], ranges[1:]):
            if b0 < a1:
                raise NotImplementedError("conductivity gradients need meshes with disjoint vertices")
        # device (triangles-first) row ranges
        for M in self.meshes:
            M["vrows"] = None if M["isolated"] else (M["vert_first"] + nt, M["vert_end"] + nt)
            M["trows"] = None if (M["current_barrier"] or M["isolated"]) else \
                (M["tri_first"] - nv, M["tri_first"] - nv + M["ntris"])
        self.parts = {}
        for k, M in enumerate(self.meshes):
            P = self.parts.setdefault(M["part"], dict(outer=[], pivot=None, pivot_mesh=None))
            if M["outermost"]:
                P["outer"].append(k)
            if M["deflation_pivot"] >= 0:
                P["pivot"], P["pivot_mesh"] = M["deflation_pivot"] + nt, k
        self.nt = nt

    def conductivities(self):
        return self._sigma.copy()

    def _apply_sigma(self, sigma):
        sigma = np.asarray(sigma, dtype=np.float64)
        if sigma.shape != self._sigma.shape:
            raise ValueError("expected %d conductivities" % len(self._sigma))
        # compare with the geometry itself: another model may share it
        for d, (old, new) in enumerate(zip(self.geom.conductivities(), sigma)):
            if new != old:
                if old == 0.0:
                    raise ValueError("the conductivity of non-conductive domain %d cannot change" % d)
                self.geom.set_conductivity(d, new)
        self._sigma = sigma.copy()
        self.pairs = self.geom.mesh_pairs()

    def gain(self, positions, moments, conductivity=None):
        """G (n_electrodes x n_dipoles) for dipoles at positions (n x 3) with
        moments (n x 3); conductivity (n_domains) defaults to the
        geometry's."""
        torch = _t()
        if conductivity is None:
            conductivity = torch.tensor(self._sigma, dtype=torch.float64)
        return _GainEEG.apply(positions, moments, conductivity, self)

    # -- forward / backward pieces --------------------------------------------
    def _forward(self, dip, sigma):
        torch = _t()
        g = self.geom
        n, ne, nd = g.size, self.ne, dip.shape[0]
        self._apply_sigma(sigma)
        doms = (C.c_int32 * max(nd, 1))()
        check(lib().omh_locate_domains(g._h, _ptr(dip), 6, nd, doms, _stream()))
        H = _head_mat_device(g, Integrator(self.order), far_field=self.precision != "fp64")
        Hkeep = H.clone() if self.precision == "fp64" else H   # the FP64 factor overwrites H
        D = _dev_f64(n * max(nd, 1))
        check(lib().omh_dip_source_mat(g._h, _ptr(dip), nd, doms, self.order, 0, 0.0, _ptr(D), n,
                                       ORDER_TRIANGLES_FIRST, _stream()))
        F = factor(H, n, self.precision)
        Y = self._E.clone()
        F.solve(Y, ne, max_iter=self.max_refine, tol=self.refine_tol)
        G = _dev_f64(ne * max(nd, 1))
        _gemm(1, 0, ne, nd, n, Y, n, D, n, G, ne)
        state = dict(H=Hkeep, F=F, Y=Y, D=D, dip=dip, doms=doms, sigma=self._sigma.copy(), nd=nd)
        return G[: ne * nd].view(nd, ne).t().contiguous(), state

    def _backward(self, st, gout, need_dip, need_sigma):
        torch = _t()
        g = self.geom
        n, ne, nd = g.size, self.ne, st["nd"]
        Gb = gout.to(device="cuda", dtype=torch.float64).t().contiguous().view(-1)   # column-major ne x nd
        W = _dev_f64(n * nd)
        _gemm(0, 0, n, nd, ne, st["Y"], n, Gb, ne, W, n)
        gdip = gsig = None
        if need_dip:
            gdip = torch.empty((nd, 6), dtype=torch.float64, device="cuda")
            check(lib().omh_dip_source_adjoint(g._h, _ptr(st["dip"]), nd, st["doms"], self.order, _ptr(W), n,
                                               ORDER_TRIANGLES_FIRST, _ptr(gdip), _stream()))
        if need_sigma:
            Z = _dev_f64(n * ne)
            _gemm(0, 1, n, ne, nd, st["D"], n, Gb, ne, Z, n)
            st["F"].solve(Z, ne, max_iter=self.max_refine, tol=self.refine_tol)
            gsig = self._sigma_gradient(st, W, Z)
        return gdip, gsig

    def _sigma_gradient(self, st, W, Z):
        torch = _t()
        n, ne, nd, nt = self.geom.size, self.ne, st["nd"], self.nt
        H, Y, D, sig = st["H"], st["Y"], st["D"], st["sigma"]
        Yv, Zv = Y.view(ne, n), Z.view(ne, n)
        grad = np.zeros(len(sig))

        def q(R, Cr):
            """<Y_R, H_RC Z_C> summed over the electrode columns."""
            (r0, r1), (c0, c1) = R, Cr
            T = _dev_f64((r1 - r0) * ne)
            _gemm(0, 0, r1 - r0, ne, c1 - c0, H, n, Z, n, T, r1 - r0, a_off=r0 + c0 * n, b_off=c0)
            return float(torch.sum(Yv[:, r0:r1] * T.view(ne, r1 - r0)))

        def ones(R):
            r0, r1 = R
            return float(torch.dot(Yv[:, r0:r1].sum(1), Zv[:, r0:r1].sum(1)))

        # deflation constants: H[i,i] = c_N N[i,i] + c with c = c_N N[i,i] / nbv
        coef = {}
        for p, P in self.parts.items():
            if P["pivot"] is None:
                continue
            nbv = sum(self.meshes[m]["nverts"] for m in P["outer"])
            i = P["pivot"]
            coef[p] = float(H[i + i * n]) / (nbv + 1)
        pair_of = {}
        for pr in self.pairs:
            m1, m2 = pr["meshes"]
            pair_of[(m1, m2)] = pr
            M1, M2 = self.meshes[m1], self.meshes[m2]
            live = [d for d in pr["domains"] if sig[d] > 0]
            if not live:
                continue
            sum_inv = sum(1.0 / sig[d] for d in pr["domains"] if sig[d] > 0)
            sum_sig = sum(sig[d] for d in pr["domains"])
            qs = 0.0
            if M1["trows"] is not None and M2["trows"] is not None:
                qs = q(M1["trows"], M2["trows"])
                if m1 != m2:
                    qs += q(M2["trows"], M1["trows"])
            qn = q(M1["vrows"], M2["vrows"])
            if m1 != m2:
                qn += q(M2["vrows"], M1["vrows"])
            elif M1["outermost"] and M1["part"] in coef:
                qn -= coef[M1["part"]] * ones(M1["vrows"])
            for d in live:
                grad[d] -= qs * (-1.0 / sig[d] ** 2) / sum_inv + qn / sum_sig
        for p, P in self.parts.items():
            if p not in coef:
                continue
            pr = pair_of.get((P["pivot_mesh"], P["pivot_mesh"]))
            if pr is None:
                continue
            sum_sig = sum(sig[d] for d in pr["domains"])
            term = coef[p] * sum(ones(self.meshes[m]["vrows"]) for m in P["outer"])
            for d in pr["domains"]:
                if sig[d] > 0:
                    grad[d] -= term / sum_sig
        # 1/sigma of the dipole's domain on the potential (triangle) rows of D
        if nt > 0 and nd > 0:
            r = (W.view(nd, n)[:, :nt] * D[: n * nd].view(nd, n)[:, :nt]).sum(1).cpu().numpy()
            doms = np.frombuffer(st["doms"], dtype=np.int32, count=nd)
            for s in range(nd):
                d = doms[s]
                if sig[d] > 0:
                    grad[d] -= r[s] / sig[d]
        return grad


def _make_function():
    torch = _t()

    class GainEEGFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, positions, moments, conductivity, model):
            pos = positions.detach().to(device="cuda", dtype=torch.float64)
            mom = moments.detach().to(device="cuda", dtype=torch.float64)
            if pos.ndim != 2 or pos.shape[1] != 3 or mom.shape != pos.shape:
                raise ValueError("positions and moments must both be n x 3")
            dip = torch.cat([pos, mom], 1).contiguous()
            sigma = conductivity.detach().cpu().numpy().astype(np.float64)
            G, st = model._forward(dip, sigma)
            ctx.model, ctx.state = model, st
            ctx.devices = (positions.device, moments.device, conductivity.device)
            ctx.dtypes = (positions.dtype, moments.dtype, conductivity.dtype)
            return G

        @staticmethod
        def backward(ctx, gout):
            need_dip = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
            gdip, gsig = ctx.model._backward(ctx.state, gout, need_dip, ctx.needs_input_grad[2])
            ctx.state = None
            gp = gm = gs = None
            if gdip is not None:
                gp = gdip[:, :3].to(device=ctx.devices[0], dtype=ctx.dtypes[0])
                gm = gdip[:, 3:].to(device=ctx.devices[1], dtype=ctx.dtypes[1])
            if gsig is not None:
                gs = torch.tensor(gsig, dtype=ctx.dtypes[2], device=ctx.devices[2])
            return gp, gm, gs, None

    return GainEEGFunction


class _GainEEG:
    """torch.autograd.Function built on first use (torch is imported lazily)."""
    fn = None

    @classmethod
    def apply(cls, *args):
        if cls.fn is None:
            cls.fn = _make_function()
        return cls.fn.apply(*args)
