"""Magnus flow drivers of IMSRG_MAGNUS.f90 on device-resident operators.

``Flow`` owns the workspaces that BCH_EXPAND / MAGNUS_EXPAND allocate per
call in the reference (INT1/INT2/AD/w1/w2, cross-coupled buffers, WCC) and
exposes the reference's routines:

    BCH_EXPAND(HS, G, H)          IMSRG_MAGNUS.f90:149-234
    MAGNUS_EXPAND(DG, G)          IMSRG_MAGNUS.f90:620-691
    euler_step(G, DG, s, ds)      IMSRG_MAGNUS.f90:754-764
    magnus_decouple(HS, G)        IMSRG_MAGNUS.f90:15-128 (white generator)

Tolerances default to the reference's single-precision literal 1e-6
(``real(8), parameter :: conv = 1e-6``), i.e. 9.999999974752427e-07.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .basis import Basis
from .operators import SqOp, _stream

REF_TOL = float(np.float32(1e-6))   # the reference's 1e-6 default-real literal


class Flow:
    def __init__(self, basis: Basis, graphs: bool = False, generator: str = "white"):
        self.basis = basis
        h = C.c_void_p()
        _lib.check(_lib.lib.imsrg_flow_create(basis.handle, C.byref(h)))
        self.handle = h
        self.set_graphs(graphs)
        self.set_generator(generator)

    def set_generator(self, name: str) -> None:
        """the generator magnus_decouple builds: "white" (build_gs_white),
        "atan" (build_gs_atan) or "w2" (build_gs_w2, the reference main's)"""
        _lib.check(_lib.lib.imsrg_flow_set_generator(self.handle, _lib.GENERATOR[name]))
        self.generator = name

    def build_gs_w2(self, H: SqOp, ETA: SqOp) -> None:
        """build_gs_w2, generators.f90:109-160 (uses this flow's workspace)"""
        ETA.herm = -1
        _lib.check(_lib.lib.imsrg_build_gs_w2(self.handle, H.ptr, ETA.ptr, _stream()))

    def transform_sq_op_BCH(self, Op: SqOp, G: SqOp, conv: float = REF_TOL, maxit: int = 40) -> int:
        """transform_sq_op_BCH, IMSRG_MAGNUS.f90:131-146 (scalar operators):
        Op <- exp(G) Op exp(-G) in place; returns BCH_EXPAND's exit index"""
        nt = C.c_int()
        _lib.check(_lib.lib.imsrg_transform_sq_op_bch(self.handle, Op.ptr, Op.herm, G.ptr, float(conv), int(maxit),
                                                      C.byref(nt), _stream()))
        return nt.value

    def dHds(self, HS: SqOp, DH: SqOp) -> None:
        """dHds, IMSRG_ODE.f90:416-451: DH = [eta(HS), HS] with this flow's
        generator (E0 included)"""
        DH.herm = 1
        _lib.check(_lib.lib.imsrg_dhds(self.handle, HS.ptr, DH.ptr, _stream()))

    def set_graphs(self, on: bool) -> None:
        _lib.check(_lib.lib.imsrg_flow_set_graphs(self.handle, 1 if on else 0))
        self.graphs = bool(on)

    def BCH_EXPAND(self, HS: SqOp, G: SqOp, H: SqOp, conv: float = REF_TOL, maxit: int = 40):
        """HS = exp(G) H exp(-G); returns (loop index of exit, ||term||*coef per term)"""
        nt = C.c_int()
        adv = (C.c_double * maxit)()
        _lib.check(_lib.lib.imsrg_bch_expand(self.handle, HS.ptr, G.ptr, G.herm, H.ptr, H.herm, float(conv),
                                             int(maxit), C.byref(nt), adv, _stream()))
        HS.herm = H.herm
        return nt.value, np.array(adv[: max(nt.value - 1, 0)])

    def MAGNUS_EXPAND(self, DG: SqOp, G: SqOp, conv: float = REF_TOL) -> int:
        """DG (holding eta) <- dOmega/ds; returns the number of nested commutators"""
        nc = C.c_int()
        _lib.check(_lib.lib.imsrg_magnus_expand(self.handle, DG.ptr, DG.herm, G.ptr, G.herm, float(conv),
                                                C.byref(nc), _stream()))
        return nc.value

    def euler_step(self, G: SqOp, DG: SqOp, s: float, ds: float) -> float:
        _lib.check(_lib.lib.imsrg_euler_step(self.basis.handle, G.ptr, DG.ptr, float(ds), _stream()))
        return s + ds

    def magnus_step(self, HS: SqOp, G: SqOp, DG: SqOp, H: SqOp, ds: float = 0.5, bch_conv: float = REF_TOL,
                    mag_conv: float = REF_TOL):
        """one iteration of magnus_decouple's loop; returns (E0, mbpt2, n_magnus, bch_index)"""
        out = (C.c_double * 4)()
        _lib.check(_lib.lib.imsrg_magnus_step(self.handle, HS.ptr, G.ptr, DG.ptr, H.ptr, float(ds),
                                              float(bch_conv), float(mag_conv), out, _stream()))
        return out[0], out[1], int(out[2]), int(out[3])

    def magnus_decouple(self, HS: SqOp, G: SqOp, ds: float = 0.5, bch_conv: float = REF_TOL,
                        mag_conv: float = REF_TOL, crit_mode: int = 0, crit_tol: float = REF_TOL,
                        maxsteps: int = 200):
        """flow HS (bare, normal ordered) to decoupling; G receives Omega.
        Returns the per-step log rows (s, E0, E0+mbpt2, nested commutators)."""
        log = (C.c_double * (4 * maxsteps))()
        st = C.c_int()
        _lib.check(_lib.lib.imsrg_magnus_decouple(self.handle, HS.ptr, G.ptr, float(ds), float(bch_conv),
                                                  float(mag_conv), int(crit_mode), float(crit_tol),
                                                  int(maxsteps), log, C.byref(st), _stream()))
        G.herm = -1
        return np.array(log[: 4 * st.value]).reshape(st.value, 4)

    def magnus_decouple_chk(self, HS: SqOp, G: SqOp, prefix: str, chk: int = 0, ds: float = 0.5,
                            bch_conv: float = REF_TOL, mag_conv: float = REF_TOL, crit_mode: int = 0,
                            crit_tol: float = REF_TOL, maxsteps: int = 200):
        """magnus_decouple with the reference's Omega checkpoints
        (IMSRG_MAGNUS.f90:64-110): restarts from the newest
        <prefix>_s<s>.bin if there is one, writes Omega every chk steps
        (0: the reference's rule).  Returns (log rows of the steps run here,
        total steps, restart s or -1)."""
        log = (C.c_double * (4 * maxsteps))()
        st = C.c_int()
        sx = C.c_double()
        _lib.check(_lib.lib.imsrg_magnus_decouple_chk(self.handle, HS.ptr, G.ptr, float(ds), float(bch_conv),
                                                      float(mag_conv), int(crit_mode), float(crit_tol),
                                                      int(maxsteps), str(prefix).encode(), int(chk), log,
                                                      C.byref(st), C.byref(sx), _stream()))
        G.herm = -1
        n0 = int(round(sx.value / ds)) if sx.value >= 0 else 0
        nrun = st.value - n0
        return np.array(log[: 4 * nrun]).reshape(nrun, 4), st.value, sx.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None and getattr(_lib, "lib", None) is not None:
            try:
                _lib.lib.imsrg_flow_destroy(h)
            except Exception:
                pass
            self.handle = None
