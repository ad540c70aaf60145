"""The reference's commutator module (commutators.f90) and operator algebra
(basic_IMSRG.f90), same routine names and argument meaning, running in
libimsrg_hip.so.

Assign / accumulate semantics follow the reference: commutator_111 and
commutator_122 overwrite RES's one- and two-body parts, the other terms add
to RES; commutator_221 needs the w1 / w2 that commutator_222_pp_hh fills;
commutator_110 / commutator_220 return numbers.  RES.herm must be set by the
caller (it drives the symmetric fills).
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from .basis import Basis
from .operators import CCMat, CCWork, SqOp, _stream

L_ = _lib.lib


def _b(op) -> Basis:
    return op.basis


def commutator_110(L: SqOp, R: SqOp) -> float:
    """commutators.f90:10-33"""
    v = C.c_double()
    _lib.check(L_.imsrg_commutator_110(_b(L).handle, L.ptr, L.herm, R.ptr, R.herm, C.byref(v), _stream()))
    return v.value


def commutator_220(L: SqOp, R: SqOp) -> float:
    """commutators.f90:399-432"""
    v = C.c_double()
    _lib.check(L_.imsrg_commutator_220(_b(L).handle, L.ptr, L.herm, R.ptr, R.herm, C.byref(v), _stream()))
    return v.value


def commutator_e0(L: SqOp, R: SqOp, RES: SqOp) -> None:
    """RES%E0 = commutator_110 + commutator_220 (IMSRG_MAGNUS.f90:210), on device"""
    _lib.check(L_.imsrg_commutator_e0(_b(L).handle, L.ptr, L.herm, R.ptr, R.herm, RES.ptr, _stream()))


def commutator_111(L: SqOp, R: SqOp, RES: SqOp) -> None:
    """commutators.f90:36-80"""
    _lib.check(L_.imsrg_commutator_111(_b(L).handle, L.ptr, L.herm, R.ptr, R.herm, RES.ptr, RES.herm, _stream()))


def commutator_121(L: SqOp, R: SqOp, RES: SqOp) -> None:
    """commutators.f90:83-280"""
    _lib.check(L_.imsrg_commutator_121(_b(L).handle, L.ptr, L.herm, R.ptr, R.herm, RES.ptr, RES.herm, _stream()))


def commutator_122(L: SqOp, R: SqOp, RES: SqOp) -> None:
    """commutators.f90:283-396"""
    _lib.check(L_.imsrg_commutator_122(_b(L).handle, L.ptr, L.herm, R.ptr, R.herm, RES.ptr, RES.herm, _stream()))


def commutator_222_pp_hh(L: SqOp, R: SqOp, RES: SqOp, w1: SqOp, w2: SqOp) -> None:
    """commutators.f90:581-768"""
    _lib.check(L_.imsrg_commutator_222_pp_hh(_b(L).handle, L.ptr, L.herm, R.ptr, R.herm, RES.ptr, RES.herm,
                                             w1.ptr, w2.ptr, _stream()))


def commutator_221(L: SqOp, R: SqOp, RES: SqOp, w1: SqOp, w2: SqOp) -> None:
    """commutators.f90:435-578"""
    _lib.check(L_.imsrg_commutator_221(_b(L).handle, L.ptr, L.herm, R.ptr, R.herm, RES.ptr, RES.herm,
                                       w1.ptr, w2.ptr, _stream()))


def calculate_cross_coupled(HS: SqOp, CCME: CCMat) -> None:
    """cross_coupled.f90:1190-1300"""
    CCME.herm = HS.herm
    _lib.check(L_.imsrg_calculate_cross_coupled(_b(HS).handle, HS.ptr, HS.herm, CCME.ptr, _stream()))


def commutator_222_ph(LCC: CCMat, RCC: CCMat, RES: SqOp, WCC: CCWork) -> None:
    """commutators.f90:771-925"""
    _lib.check(L_.imsrg_commutator_222_ph(_b(RES).handle, LCC.ptr, LCC.herm, RCC.ptr, RCC.herm, RES.ptr,
                                          RES.herm, WCC.ptr, _stream()))


def init_ph_mat(OP: SqOp) -> CCMat:
    """allocate_CCMAT, cross_coupled.f90:61-177"""
    return CCMat(OP.basis, OP.herm, OP.data.device)


def duplicate_ph_mat(C1: CCMat) -> CCMat:
    """dup_CC, cross_coupled.f90:863-914"""
    return CCMat(C1.basis, C1.herm, C1.data.device)


def init_ph_wkspc(CCOP: CCMat) -> CCWork:
    """allocate_CC_wkspc, cross_coupled.f90:1075-1096"""
    return CCWork(CCOP.basis, CCOP.data.device)


# ------------------------------------------------------------ operator algebra
def duplicate_sq_op(H: SqOp) -> SqOp:
    """duplicate_sq_op, basic_IMSRG.f90:2797-2860 (zeroed, herm = 1)"""
    return SqOp.zeros(H.basis, 1, H.data.device)


def copy_sq_op(H: SqOp, op: SqOp) -> None:
    """copy_sq_op, basic_IMSRG.f90:3005-3032 (copies herm too)"""
    op.data.copy_(H.data)
    op.herm = H.herm


def clear_sq_op(C_: SqOp) -> None:
    """clear_sq_op, basic_IMSRG.f90:3638-3663"""
    C_.data.zero_()


def add_sq_op(A: SqOp, ax: float, B: SqOp, bx: float, C_: SqOp) -> None:
    """add_sq_op, basic_IMSRG.f90:3551-3583: C = ax*A + bx*B"""
    _lib.check(L_.imsrg_add_sq_op(_b(A).handle, A.ptr, float(ax), B.ptr, float(bx), C_.ptr, _stream()))


def mat_frob_norm(op: SqOp) -> float:
    """mat_frob_norm, basic_IMSRG.f90:4056-4087"""
    v = C.c_double()
    _lib.check(L_.imsrg_mat_frob_norm(_b(op).handle, op.ptr, C.byref(v), _stream()))
    return v.value


def build_gs_white(H: SqOp, ETA: SqOp) -> None:
    """build_gs_white, generators.f90:4-106"""
    ETA.herm = -1
    _lib.check(L_.imsrg_build_gs_white(_b(H).handle, H.ptr, ETA.ptr, _stream()))


def build_gs_atan(H: SqOp, ETA: SqOp) -> None:
    """build_gs_atan, generators.f90:163-269"""
    ETA.herm = -1
    _lib.check(L_.imsrg_build_gs_atan(_b(H).handle, H.ptr, ETA.ptr, _stream()))


def build_gs_atanxx(Hold: SqOp, H: SqOp, ETA: SqOp) -> None:
    """build_gs_atanxx, generators.f90:271-374 (denominators of Hold)"""
    ETA.herm = -1
    _lib.check(L_.imsrg_build_gs_atanxx(_b(H).handle, Hold.ptr, H.ptr, ETA.ptr, _stream()))


def mbpt2(H: SqOp) -> float:
    """mbpt2, basic_IMSRG.f90:4165-4199"""
    v = C.c_double()
    _lib.check(L_.imsrg_mbpt2(_b(H).handle, H.ptr, C.byref(v), _stream()))
    return v.value


def normal_order(H: SqOp) -> None:
    """normal_order, basic_IMSRG.f90:2419-2550 (in place)"""
    _lib.check(L_.imsrg_normal_order(_b(H).handle, H.ptr, _stream()))


def commutator(L: SqOp, R: SqOp, RES: SqOp, w1: SqOp | None = None, w2: SqOp | None = None) -> None:
    """RES = [L, R] with the term sequence of BCH_EXPAND (IMSRG_MAGNUS.f90:207-219)"""
    b = L.basis
    w1 = w1 or SqOp.zeros(b, 1, L.data.device)
    w2 = w2 or SqOp.zeros(b, 1, L.data.device)
    LCC, RCC = init_ph_mat(L), init_ph_mat(R)
    WCC = init_ph_wkspc(LCC)
    calculate_cross_coupled(R, RCC)
    calculate_cross_coupled(L, LCC)
    commutator_e0(L, R, RES)
    commutator_111(L, R, RES)
    commutator_121(L, R, RES)
    commutator_122(L, R, RES)
    commutator_222_pp_hh(L, R, RES, w1, w2)
    commutator_221(L, R, RES, w1, w2)
    commutator_222_ph(LCC, RCC, RES, WCC)
