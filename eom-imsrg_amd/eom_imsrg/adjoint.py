"""Reverse-mode derivatives of the Magnus flow (the reference has none).

The C-ABI's hand-written transposes (imsrg_*_vjp, imsrg_flow_grad) give
dE0/d(input matrix elements) for the flow of magnus_decouple
(IMSRG_MAGNUS.f90:15-128, white generator).  ``flow_energy`` exposes it to
PyTorch: a ``torch.autograd.Function`` whose forward runs the flow from the
vacuum operator T (normal ordered on the device) and whose backward runs the
reverse sweep with Omega checkpointed every ``ckpt`` steps (``ckpt=0``: every
step kept in device memory, no recomputation).

Gradients are with respect to the flat operator arrays; they are exact for
perturbations that keep herm symmetry and the zero Pauli-forbidden (a,a)
odd-J elements (``symmetric_gradient`` folds a gradient onto the stored upper
elements, the derivative with respect to a hermitian matrix element).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .basis import Basis, SQUARE
from .magnus import Flow, REF_TOL
from .operators import SqOp, _stream


def commutator_vjp(L: SqOp, R: SqOp, RESb: torch.Tensor, RESh: int, Lb: torch.Tensor, Rb: torch.Tensor) -> None:
    """Lb += d<RESb,[L,R]>/dL, Rb += d<RESb,[L,R]>/dR"""
    _lib.check(_lib.lib.imsrg_commutator_vjp(L.basis.handle, L.ptr, L.herm, R.ptr, R.herm, RESb.data_ptr(), RESh,
                                             Lb.data_ptr(), Rb.data_ptr(), _stream()))


def bch_vjp(flow: Flow, G: SqOp, H: SqOp, HSb: torch.Tensor, Gb: torch.Tensor, Hb: torch.Tensor,
            conv: float = REF_TOL, maxit: int = 40) -> int:
    nt = C.c_int()
    _lib.check(_lib.lib.imsrg_bch_vjp(flow.handle, G.ptr, H.ptr, H.herm, float(conv), int(maxit), HSb.data_ptr(),
                                      Gb.data_ptr(), Hb.data_ptr(), C.byref(nt), _stream()))
    return nt.value


def magnus_vjp(flow: Flow, ETA: SqOp, G: SqOp, DGb: torch.Tensor, ETAb: torch.Tensor, Gb: torch.Tensor,
               conv: float = REF_TOL) -> int:
    nc = C.c_int()
    _lib.check(_lib.lib.imsrg_magnus_vjp(flow.handle, ETA.ptr, G.ptr, float(conv), DGb.data_ptr(), ETAb.data_ptr(),
                                         Gb.data_ptr(), C.byref(nc), _stream()))
    return nc.value


def white_vjp(H: SqOp, ETAb: torch.Tensor, Hb: torch.Tensor) -> None:
    _lib.check(_lib.lib.imsrg_white_vjp(H.basis.handle, H.ptr, ETAb.data_ptr(), Hb.data_ptr(), _stream()))


def generator_vjp(flow: Flow, H: SqOp, ETAb: torch.Tensor, Hb: torch.Tensor) -> None:
    """transpose of the flow's generator (white, atan or w2): Hb += d<ETAb, eta(H)>/dH"""
    _lib.check(_lib.lib.imsrg_generator_vjp(flow.handle, H.ptr, ETAb.data_ptr(), Hb.data_ptr(), _stream()))


def normal_order_vjp(basis: Basis, Hb: torch.Tensor, Tb: torch.Tensor) -> None:
    _lib.check(_lib.lib.imsrg_normal_order_vjp(basis.handle, Hb.data_ptr(), Tb.data_ptr(), _stream()))


def flow_grad(flow: Flow, T: torch.Tensor, ds: float = 0.5, bch_conv: float = REF_TOL, mag_conv: float = REF_TOL,
              crit_tol: float = REF_TOL, maxsteps: int = 200, ckpt: int = 4):
    """(E0, dE0/dT, steps) of the flow from normal_order(T)"""
    Tb = torch.empty_like(T)
    e0 = C.c_double()
    st = C.c_int()
    _lib.check(_lib.lib.imsrg_flow_grad(flow.handle, T.data_ptr(), float(ds), float(bch_conv), float(mag_conv),
                                        float(crit_tol), int(maxsteps), int(ckpt), Tb.data_ptr(), C.byref(e0),
                                        C.byref(st), _stream()))
    return e0.value, Tb, st.value


def flow_forward(flow: Flow, T: torch.Tensor, ds: float = 0.5, bch_conv: float = REF_TOL,
                 mag_conv: float = REF_TOL, crit_tol: float = REF_TOL, maxsteps: int = 200, ckpt: int = 4):
    """run the flow from normal_order(T) keeping Omega every `ckpt` steps in
    `flow` (ckpt 0: every step's states and terms, so that flow_backward
    recomputes nothing); returns (E0, steps)"""
    e0 = C.c_double()
    st = C.c_int()
    _lib.check(_lib.lib.imsrg_flow_forward(flow.handle, T.data_ptr(), float(ds), float(bch_conv), float(mag_conv),
                                           float(crit_tol), int(maxsteps), int(ckpt), C.byref(e0), C.byref(st),
                                           _stream()))
    return e0.value, st.value


def flow_backward(flow: Flow, Tb: torch.Tensor) -> None:
    """dE0/dT of the last flow_forward on `flow` into Tb"""
    _lib.check(_lib.lib.imsrg_flow_backward(flow.handle, Tb.data_ptr(), _stream()))


class _FlowEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, T, flow, ds, tol, ckpt):
        Tc = T.detach().contiguous()
        e0, _ = flow_forward(flow, Tc, ds, tol, tol, tol, ckpt=ckpt)
        ctx.flow = flow
        return torch.tensor(e0, dtype=torch.float64, device=T.device)

    @staticmethod
    def backward(ctx, gE):
        Tb = torch.empty(ctx.flow.basis.oplen, dtype=torch.float64, device=gE.device)
        flow_backward(ctx.flow, Tb)
        return Tb * gE, None, None, None, None


def flow_energy(T: torch.Tensor, flow: Flow, ds: float = 0.5, tol: float = REF_TOL, ckpt: int = 4) -> torch.Tensor:
    """E0 after magnus_decouple from the vacuum operator T (CUDA float64,
    basis.oplen elements), differentiable with respect to T.  The flow object
    holds the checkpoints of the forward sweep until the next call, so call
    backward before running another flow_energy on the same `flow`."""
    return _FlowEnergy.apply(T, flow, ds, tol, ckpt)


def symmetric_gradient(basis: Basis, g: np.ndarray, herm: int) -> np.ndarray:
    """fold a flat gradient onto the independent elements of an operator
    with herm symmetry: square blocks (one-body hh / pp and gam1, 4, 5) keep
    g[up] + herm * g[lo] on the upper triangle and zero below (the diagonal
    is zeroed for herm = -1); rows / columns of Pauli-forbidden pairs (a,a)
    at odd J and one-body elements between different (l, j, tz) blocks are
    zeroed.  The result is the derivative with respect to each independent
    (anti)hermitian matrix element."""
    b = basis
    out = np.array(g, dtype=np.float64, copy=True)

    def fold_sq(off, n):
        X = out[off: off + n * n].reshape(n, n, order="F")
        Y = np.triu(X + herm * X.T) if herm == 1 else np.triu(X + herm * X.T, 1)
        if herm == 1:
            np.fill_diagonal(Y, np.diag(X))
        out[off: off + n * n] = Y.reshape(-1, order="F")

    A, P = b.nhole, b.npart
    key = lambda a: (int(b.l[a]), int(b.jj[a]), int(b.tz[a]))
    same_hh = np.array([[key(x) == key(y) for y in b.holes] for x in b.holes], bool)
    same_pp = np.array([[key(x) == key(y) for y in b.parts] for x in b.parts], bool)
    same_ph = np.array([[key(x) == key(y) for y in b.holes] for x in b.parts], bool)
    fold_sq(b.off_fhh, A)
    out[b.off_fhh: b.off_fhh + A * A] *= same_hh.reshape(-1, order="F")
    out[b.off_fph: b.off_fph + P * A] *= same_ph.reshape(-1, order="F")
    fold_sq(b.off_fpp, P)
    out[b.off_fpp: b.off_fpp + P * P] *= same_pp.reshape(-1, order="F")
    rows_of = (0, 0, 0, 1, 2, 1)
    cols_of = (0, 1, 2, 1, 2, 2)
    for q in range(b.nblocks):
        odd = (int(b.jtp[q, 0]) // 2) % 2 == 1
        for g_ in range(6):
            n1, n2 = b.block_shape(q, g_)
            if n1 * n2 == 0:
                continue
            o = int(b.goff[q, g_])
            if SQUARE[g_]:
                fold_sq(o, n1)
            if odd:
                X = out[o: o + n1 * n2].reshape(n1, n2, order="F")
                pr, pc = b.pairs[q][rows_of[g_]], b.pairs[q][cols_of[g_]]
                X[pr[:, 0] == pr[:, 1], :] = 0.0
                X[:, pc[:, 0] == pc[:, 1]] = 0.0
                out[o: o + n1 * n2] = X.reshape(-1, order="F")
    return out
