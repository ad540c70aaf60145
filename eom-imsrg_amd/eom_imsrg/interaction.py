"""A physical Hamiltonian on the basis: the harmonic-oscillator one-body part
(calculate_h0_harm_osc, basic_IMSRG.f90:2093-2191) and the two-body matrix
elements of an me2j interaction file (read_me2j_interaction,
interaction_IO.f90:282-557), as the reference's main program sets them up
before normal ordering (main_IMSRG.f90:121-140).

The scalars the reference keeps on the Hamiltonian (hospace, lawson_beta,
com_hw, Aprot + Aneut) are arguments; ``htype`` is the reference's
Hamiltonian type (1: T - Tcm + V, 2: T + Uho + V, 3: T + V).  The vacuum
Hamiltonian is then normal ordered either in the HO basis
(``commutators.normal_order``) or in the Hartree-Fock basis
(``hartree_fock``, calc_HF of HF_mod.f90:14-105).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .operators import SqOp, _stream
from .orbits import hk_orbits


def h0_harm_osc(H: SqOp, hw: float, htype: int = 1, mass: int = 16, lawson_beta: float = 0.0,
                com_hw: float | None = None) -> SqOp:
    """fhh, fph, fpp of H <- the HO one-body Hamiltonian of type htype"""
    _lib.check(_lib.lib.imsrg_h0_harm_osc(H.basis.handle, float(hw), int(htype), float(lawson_beta),
                                          float(hw if com_hw is None else com_hw), int(mass), H.ptr, _stream()))
    return H


def file_orbits(file_emax: int, file_lmax: int | None = None) -> dict:
    """orbit table of the .sps file an me2j file of that eMax is written on"""
    return hk_orbits(file_emax, file_lmax)


def read_me2j(H: SqOp, path: str, file_emax: int | None = None, hw: float = 20.0, htype: int = 1, mass: int = 16,
              lawson_beta: float = 0.0, com_hw: float | None = None, orbits: dict | None = None) -> SqOp:
    """two-body blocks of H <- the me2j file's matrix elements (proton-neutron
    form, centre-of-mass and Lawson terms added); ``orbits`` (n, l, jj, tz)
    is the file's orbit table, by default the hk table of ``file_emax``"""
    if orbits is None:
        if file_emax is None:
            raise ValueError("give file_emax or the file's orbit table")
        orbits = file_orbits(file_emax)
    arr = [np.ascontiguousarray(orbits[k], np.int32) for k in ("n", "l", "jj", "tz")]
    ptrs = [a.ctypes.data_as(C.POINTER(C.c_int)) for a in arr]
    _lib.check(_lib.lib.imsrg_read_me2j(H.basis.handle, str(path).encode(), len(arr[0]), *ptrs, float(hw), int(htype),
                                        float(lawson_beta), float(hw if com_hw is None else com_hw), int(mass),
                                        H.ptr, _stream()))
    return H


def hartree_fock(H: SqOp, maxit: int = 200):
    """H (vacuum (t, V), HO basis) <- normal ordered in the Hartree-Fock basis;
    returns (e_HF, iterations, HF orbits per (tz, l, j) block, flat
    column-major)"""
    e = C.c_double()
    it = C.c_int()
    n = int(_lib.lib.imsrg_basis_size(H.basis.handle, _lib.SIZE["SPMAT"]))
    D = np.zeros(max(n, 1))
    _lib.check(_lib.lib.imsrg_hartree_fock(H.basis.handle, H.ptr, int(maxit), C.byref(e), C.byref(it),
                                           D.ctypes.data_as(C.POINTER(C.c_double)), _stream()))
    return e.value, it.value, D[:n]
