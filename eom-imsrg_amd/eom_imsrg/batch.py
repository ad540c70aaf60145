"""Many-instance batches: the step-level drivers of IMSRG_MAGNUS.f90
(BCH_EXPAND 149-234, MAGNUS_EXPAND 620-691, magnus_decouple 15-128) for
``ninst`` independent problems on one basis at once.

Operators are CUDA float64 tensors of shape (ninst, basis.oplen); row k is
instance k's flat operator.  Every kernel launch covers all instances, and
each instance takes the BCH / Magnus / convergence exits of its own flow, so
row k of a batched result equals what a single ``Flow`` gives for row k.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .basis import Basis
from .magnus import REF_TOL
from .operators import _stream


class Batch:
    def __init__(self, basis: Basis, ninst: int, graphs: bool = True, generator: str = "white"):
        self.basis = basis
        self.ninst = int(ninst)
        h = C.c_void_p()
        _lib.check(_lib.lib.imsrg_batch_create(basis.handle, self.ninst, C.byref(h)))
        self.handle = h
        _lib.check(_lib.lib.imsrg_flow_set_graphs(h, 1 if graphs else 0))
        self.graphs = bool(graphs)
        _lib.check(_lib.lib.imsrg_flow_set_generator(h, _lib.GENERATOR[generator]))
        self.generator = generator

    def _check(self, *ts: torch.Tensor) -> None:
        for t in ts:
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
                raise ValueError("batched operators are contiguous CUDA float64 tensors")
            if tuple(t.shape) != (self.ninst, self.basis.oplen):
                raise ValueError(f"expected shape ({self.ninst}, {self.basis.oplen}), got {tuple(t.shape)}")

    def zeros(self) -> torch.Tensor:
        return torch.zeros(self.ninst, self.basis.oplen, dtype=torch.float64, device="cuda")

    def commutator(self, L: torch.Tensor, Lh: int, R: torch.Tensor, Rh: int, RES: torch.Tensor) -> None:
        """RES[k] = [L[k], R[k]] (E0 included), RES herm = -Lh*Rh"""
        self._check(L, R, RES)
        _lib.check(_lib.lib.imsrg_batch_commutator(self.handle, L.data_ptr(), int(Lh), R.data_ptr(), int(Rh),
                                                   RES.data_ptr(), _stream()))

    def BCH_EXPAND(self, HS: torch.Tensor, G: torch.Tensor, H: torch.Tensor, conv: float = REF_TOL,
                   maxit: int = 40) -> np.ndarray:
        """HS[k] = exp(G[k]) H[k] exp(-G[k]); returns the exit index per instance"""
        self._check(HS, G, H)
        nt = (C.c_int * self.ninst)()
        _lib.check(_lib.lib.imsrg_batch_bch_expand(self.handle, HS.data_ptr(), G.data_ptr(), H.data_ptr(),
                                                   float(conv), int(maxit), nt, _stream()))
        return np.array(nt[:])

    def MAGNUS_EXPAND(self, DG: torch.Tensor, G: torch.Tensor, conv: float = REF_TOL) -> np.ndarray:
        """DG[k] (holding eta) <- dOmega/ds; returns nested commutators per instance"""
        self._check(DG, G)
        nc = (C.c_int * self.ninst)()
        _lib.check(_lib.lib.imsrg_batch_magnus_expand(self.handle, DG.data_ptr(), G.data_ptr(), float(conv), nc,
                                                      _stream()))
        return np.array(nc[:])

    def magnus_decouple(self, HS: torch.Tensor, G: torch.Tensor, ds: float = 0.5, bch_conv: float = REF_TOL,
                        mag_conv: float = REF_TOL, crit_tol: float = REF_TOL, maxsteps: int = 200):
        """flow every HS[k] (normal ordered) to decoupling, G[k] (zero on entry)
        receives Omega; returns (E0 per instance, steps per instance, batched
        steps run)"""
        self._check(HS, G)
        e0 = (C.c_double * self.ninst)()
        st = (C.c_int * self.ninst)()
        nmax = C.c_int()
        _lib.check(_lib.lib.imsrg_batch_magnus_decouple(self.handle, HS.data_ptr(), G.data_ptr(), float(ds),
                                                        float(bch_conv), float(mag_conv), float(crit_tol),
                                                        int(maxsteps), e0, st, C.byref(nmax), _stream()))
        return np.array(e0[:]), np.array(st[:]), nmax.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None and getattr(_lib, "lib", None) is not None:
            try:
                _lib.lib.imsrg_flow_destroy(h)
            except Exception:
                pass
            self.handle = None
