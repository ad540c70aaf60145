"""Tensor operators and the scalar-tensor commutator on the device: the
Python mirror of TS_commutators.f90 (the routine names and argument order of
the reference) and of BCH_TENSOR / transform_sq_op_BCH (IMSRG_MAGNUS.f90:
131-146, 389-464).  Every call runs in libimsrg_hip.so.

A tensor operator of rank lambda carries ``rank = 2 lambda`` and the parity
change ``dpar = 2 * (0 or 1)``, "multiplied by 2" as the reference's sq_op
does; its flat layout is described in include/imsrg_hip.h.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .basis import Basis
from .operators import CCMat, SqOp, _stream


_DESTROY = _lib.lib.imsrg_tensor_destroy   # kept for interpreter shutdown


class TensorLayout:
    """allocate_tensor (basic_IMSRG.f90:688-951) for one (rank, dpar)"""

    def __init__(self, basis: Basis, rank: int, dpar: int = 0):
        h = C.c_void_p()
        _lib.check(_lib.lib.imsrg_tensor_create(basis.handle, int(rank), int(dpar), C.byref(h)))
        self.handle = h.value
        self.basis = basis
        self.rank, self.dpar = int(rank), int(dpar)
        sz = lambda w: int(_lib.lib.imsrg_tensor_size(self.handle, _lib.TSIZE[w]))
        self.oplen, self.nblocks, self.ncc, self.cclen, self.wlen = (sz(k) for k in ("OP", "NBLOCKS", "NCC", "CC", "W"))
        nb = self.nblocks
        info = np.zeros(7 * nb, np.int32)
        dims = np.zeros(6 * nb, np.int32)
        goff = np.zeros(9 * nb, np.int64)
        _lib.check(_lib.lib.imsrg_tensor_blocks(self.handle, info.ctypes.data_as(C.POINTER(C.c_int)),
                                                dims.ctypes.data_as(C.POINTER(C.c_int)),
                                                goff.ctypes.data_as(C.POINTER(C.c_int64))))
        self.info, self.dims, self.goff = info.reshape(nb, 7), dims.reshape(nb, 6), goff.reshape(nb, 9)

    SEA1 = (0, 0, 0, 1, 2, 1, 2, 1, 2)
    SEA2 = (0, 1, 2, 1, 2, 2, 0, 0, 1)

    def shape(self, q: int, g: int):
        return int(self.dims[q, self.SEA1[g]]), int(self.dims[q, 3 + self.SEA2[g]])

    def elements(self):
        """(q, g, ix, jx, (a, b), (c, d), 2J1, 2J2, offset) of every stored two-body element"""
        b = self.basis
        for q in range(self.nblocks):
            J1, J2, _, _, _, q1, q2 = (int(x) for x in self.info[q])
            if q2 < 0:
                continue
            for g in range(9):
                n1, n2 = self.shape(q, g)
                if n1 * n2 == 0:
                    continue
                rows, cols = b.pairs[q1][self.SEA1[g]], b.pairs[q2][self.SEA2[g]]
                for jx in range(n2):
                    for ix in range(n1):
                        yield (q, g, ix, jx, (int(rows[ix][0]), int(rows[ix][1])),
                               (int(cols[jx][0]), int(cols[jx][1])), J1, J2, int(self.goff[q, g]) + ix + jx * n1)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _DESTROY is not None:
            _DESTROY(h)
            self.handle = None


class TensorOp:
    """a tensor operator on the device (sq_op with rank > 0)"""

    def __init__(self, layout: TensorLayout, data: torch.Tensor, herm: int = 1):
        if data.dtype != torch.float64 or data.numel() != layout.oplen or not data.is_cuda:
            raise ValueError("TensorOp needs a CUDA float64 tensor of layout.oplen elements")
        self.layout = layout
        self.data = data.contiguous()
        self.herm = int(herm)

    @classmethod
    def zeros(cls, layout: TensorLayout, herm: int = 1, device="cuda") -> "TensorOp":
        return cls(layout, torch.zeros(layout.oplen, dtype=torch.float64, device=device), herm)

    @classmethod
    def from_numpy(cls, layout: TensorLayout, arr: np.ndarray, herm: int = 1, device="cuda") -> "TensorOp":
        return cls(layout, torch.from_numpy(np.ascontiguousarray(arr, np.float64)).to(device), herm)

    def numpy(self) -> np.ndarray:
        return self.data.detach().cpu().numpy()

    @property
    def ptr(self) -> int:
        return self.data.data_ptr()

    @property
    def rank(self) -> int:
        return self.layout.rank

    def duplicate(self, herm: int | None = None) -> "TensorOp":
        return TensorOp.zeros(self.layout, self.herm if herm is None else herm, self.data.device)

    def clone(self) -> "TensorOp":
        return TensorOp(self.layout, self.data.clone(), self.herm)


class TensorCC:
    """pandya_mat of a tensor (allocate_small_tensor_CCMAT, cross_coupled.f90:498-700)"""

    def __init__(self, layout: TensorLayout, device="cuda"):
        self.layout = layout
        self.data = torch.zeros(max(layout.cclen, 1), dtype=torch.float64, device=device)

    @property
    def ptr(self) -> int:
        return self.data.data_ptr()


def _h(t) -> int:
    return t.layout.handle


# ------------------------------------------------------------ the terms
def TS_commutator_111(L: SqOp, R: TensorOp, RES: TensorOp) -> None:
    """TS_commutators.f90:10-55"""
    _lib.check(_lib.lib.imsrg_ts_commutator_111(_h(R), L.ptr, L.herm, R.ptr, R.herm, RES.ptr, RES.herm, _stream()))


def TS_commutator_121(L: SqOp, R: TensorOp, RES: TensorOp) -> None:
    """TS_commutators.f90:58-303"""
    _lib.check(_lib.lib.imsrg_ts_commutator_121(_h(R), L.ptr, L.herm, R.ptr, R.herm, RES.ptr, RES.herm, _stream()))


def TS_commutator_211(LCC: CCMat, R: TensorOp, RES: TensorOp) -> None:
    """TS_commutators.f90:307-552 (LCC: calculate_cross_coupled of the scalar L)"""
    _lib.check(_lib.lib.imsrg_ts_commutator_211(_h(R), LCC.ptr, LCC.herm, R.ptr, R.herm, RES.ptr, RES.herm,
                                                _stream()))


def TS_commutator_122(L: SqOp, R: TensorOp, RES: TensorOp) -> None:
    """TS_commutators.f90:555-663"""
    _lib.check(_lib.lib.imsrg_ts_commutator_122(_h(R), L.ptr, L.herm, R.ptr, R.herm, RES.ptr, RES.herm, _stream()))


def TS_commutator_212(L: SqOp, R: TensorOp, RES: TensorOp) -> None:
    """TS_commutators.f90:666-806"""
    _lib.check(_lib.lib.imsrg_ts_commutator_212(_h(R), L.ptr, L.herm, R.ptr, R.herm, RES.ptr, RES.herm, _stream()))


def TS_commutator_222_pp_hh(L: SqOp, R: TensorOp, RES: TensorOp, w1: TensorOp, w2: TensorOp) -> None:
    """TS_commutators.f90:1070-1456"""
    _lib.check(_lib.lib.imsrg_ts_commutator_222_pp_hh(_h(R), L.ptr, L.herm, R.ptr, R.herm, RES.ptr, RES.herm,
                                                      w1.ptr, w2.ptr, _stream()))


def TS_commutator_221(w1: TensorOp, w2: TensorOp, pm: int, RES: TensorOp) -> None:
    """TS_commutators.f90:809-1067"""
    _lib.check(_lib.lib.imsrg_ts_commutator_221(_h(RES), w1.ptr, w2.ptr, int(pm), RES.ptr, RES.herm, _stream()))


def calculate_single_pandya(R: TensorOp, RCC: TensorCC) -> None:
    """cross_coupled.f90:1416-1634, every block and both terms"""
    _lib.check(_lib.lib.imsrg_calculate_single_pandya(_h(R), R.ptr, R.herm, RCC.ptr, _stream()))


def TS_commutator_222_ph(LCC: CCMat, RCC: TensorCC, R: TensorOp, RES: TensorOp) -> None:
    """TS_commutators.f90:1459-1849"""
    _lib.check(_lib.lib.imsrg_ts_commutator_222_ph(_h(R), LCC.ptr, LCC.herm, RCC.ptr, R.ptr, R.herm, RES.ptr,
                                                   RES.herm, _stream()))


def TS_commutator(L: SqOp, R: TensorOp, RES: TensorOp) -> None:
    """RES = [L, R], the term sequence of BCH_TENSOR (IMSRG_MAGNUS.f90:425-446)"""
    _lib.check(_lib.lib.imsrg_ts_commutator(_h(R), L.ptr, L.herm, R.ptr, R.herm, RES.ptr, RES.herm, _stream()))


def mat_frob_norm(op: TensorOp) -> float:
    """basic_IMSRG.f90:4056-4087, rank > 0 branch"""
    v = C.c_double()
    _lib.check(_lib.lib.imsrg_tensor_frob_norm(_h(op), op.ptr, C.byref(v), _stream()))
    return v.value


def append_operator(B: TensorOp, bx: float, C_: TensorOp) -> None:
    """C = C + bx B (basic_IMSRG.f90:3582-3611)"""
    _lib.check(_lib.lib.imsrg_tensor_add(_h(C_), C_.ptr, 1.0, B.ptr, float(bx), C_.ptr, _stream()))


BCH_CONV = float(np.float32(1e-6))   # real(8), parameter :: conv = 1e-6 (IMSRG_MAGNUS.f90:392)


def BCH_TENSOR(G: SqOp, HS: TensorOp, conv: float = BCH_CONV):
    """HS <- e^G HS e^-G in place (IMSRG_MAGNUS.f90:389-464); returns (last
    term iw, advals)"""
    it = C.c_int()
    adv = (C.c_double * 30)()
    _lib.check(_lib.lib.imsrg_bch_tensor(_h(HS), G.ptr, G.herm, HS.ptr, HS.herm, float(conv), C.byref(it), adv,
                                         _stream()))
    return it.value, np.array(adv[:])


def transform_sq_op_BCH(Op, G: SqOp, flow=None, conv: float | None = None):
    """IMSRG_MAGNUS.f90:131-146: a tensor goes through BCH_TENSOR, a scalar
    through BCH_EXPAND (with `flow`, a magnus.Flow of the basis)"""
    if isinstance(Op, TensorOp):
        return BCH_TENSOR(G, Op, BCH_CONV if conv is None else conv)
    if flow is None:
        raise ValueError("a scalar operator needs a Flow for BCH_EXPAND")
    out = Op.duplicate()
    res = flow.BCH_EXPAND(out, G, Op, **({} if conv is None else {"conv": conv}))
    Op.data.copy_(out.data)
    return res
