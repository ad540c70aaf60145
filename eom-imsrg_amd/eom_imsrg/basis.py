"""Basis handle: orbit table, channels, pair lists and cross-coupled layout.

Mirrors the reference's ``type(spd)`` plus the channel structure that
``allocate_blocks`` (basic_IMSRG.f90:521-685) and ``allocate_CCMAT``
(cross_coupled.f90:61-177) attach to every operator.  The tables live in
libimsrg_hip.so; this class reads them back for host-side use (synthetic
inputs, work accounting, tests).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .orbits import hk_orbits, oxygen16_holes

SQUARE = (True, False, False, True, True, False)
ROW_CLS = (0, 0, 0, 1, 2, 1)
COL_CLS = (0, 1, 2, 1, 2, 2)


class Basis:
    def __init__(self, n, l, jj, tz, hole, e=None):
        self.n = np.ascontiguousarray(n, np.int32)
        self.l = np.ascontiguousarray(l, np.int32)
        self.jj = np.ascontiguousarray(jj, np.int32)
        self.tz = np.ascontiguousarray(tz, np.int32)
        self.hole = np.ascontiguousarray(hole, np.int32)
        self.e = np.ascontiguousarray(e if e is not None else 2 * self.n + self.l, np.int32)
        N = len(self.n)
        h = C.c_void_p()
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        _lib.check(_lib.lib.imsrg_basis_create(N, ip(self.n), ip(self.l), ip(self.jj), ip(self.tz),
                                               ip(self.hole), C.byref(h)))
        self.handle = h
        size = lambda k: int(_lib.lib.imsrg_basis_size(h, _lib.SIZE[k]))
        self.norb = size("NORB")
        self.nhole = size("NHOLE")
        self.npart = size("NPART")
        self.jtotal_max = size("JTOTAL_MAX")
        self.nblocks = size("NBLOCKS")
        self.oplen = size("OP")
        self.ncc = size("NCC")
        self.cclen = size("CC")
        self.wcclen = size("WCC")
        self.neq = size("NEQ")
        nb = self.nblocks
        self.dims = np.zeros(nb * 3, np.int32)
        self.goff = np.zeros(nb * 6, np.int64)
        self.jtp = np.zeros(nb * 3, np.int32)
        _lib.check(_lib.lib.imsrg_basis_channels(h, ip(self.dims), self.goff.ctypes.data_as(C.POINTER(C.c_int64)),
                                                 ip(self.jtp)))
        self.dims = self.dims.reshape(nb, 3)
        self.goff = self.goff.reshape(nb, 6)
        self.jtp = self.jtp.reshape(nb, 3)
        self.pairs = []
        for q in range(nb):
            row = []
            for c in range(3):
                buf = np.zeros(2 * max(int(self.dims[q, c]), 1), np.int32)
                k = _lib.lib.imsrg_basis_pairs(h, q, c, ip(buf))
                row.append(buf[: 2 * k].reshape(k, 2))
            self.pairs.append(row)
        m = self.ncc
        self.rlen = np.zeros(m, np.int32)
        self.ncol = np.zeros(m, np.int32)
        self.ccoff = np.zeros(m, np.int64)
        self.wcoff = np.zeros(m, np.int64)
        self.ccjtp = np.zeros(m * 3, np.int32)
        lp = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
        _lib.check(_lib.lib.imsrg_basis_cc(h, ip(self.rlen), ip(self.ncol), lp(self.ccoff), lp(self.wcoff),
                                           ip(self.ccjtp)))
        self.ccjtp = self.ccjtp.reshape(m, 3)
        A, P = self.nhole, self.npart
        self.off_fhh = 1
        self.off_fph = 1 + A * A
        self.off_fpp = self.off_fph + P * A
        self.holes = np.nonzero(self.hole == 1)[0]
        self.parts = np.nonzero(self.hole == 0)[0]

    @classmethod
    def oxygen16(cls, emax: int, lmax: int | None = None) -> "Basis":
        """HO basis of the configurations: hk orbits 2n+l <= emax, 16O holes."""
        o = hk_orbits(emax, lmax)
        return cls(o["n"], o["l"], o["jj"], o["tz"], oxygen16_holes(o), o["e"])

    def block_shape(self, q: int, g: int):
        return int(self.dims[q, ROW_CLS[g]]), int(self.dims[q, COL_CLS[g]])

    def set_path(self, path: int) -> None:
        _lib.check(_lib.lib.imsrg_set_path(self.handle, path))

    def set_term_streams(self, on: bool) -> None:
        """grouped commutator: independent terms on auxiliary streams (default on)"""
        _lib.check(_lib.lib.imsrg_set_term_streams(self.handle, 1 if on else 0))

    def sixj(self, j1, j2, j3, j4, j5, j6) -> float:
        return float(_lib.lib.imsrg_basis_sixj(self.handle, j1, j2, j3, j4, j5, j6))

    def e2_count(self) -> int:
        """number of stored two-body elements (all six blocks of all channels)"""
        return int(self.oplen - self.off_fpp - self.npart * self.npart)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None and getattr(_lib, "lib", None) is not None:
            try:
                _lib.lib.imsrg_basis_destroy(h)
            except Exception:
                pass
            self.handle = None
